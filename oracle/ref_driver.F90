! ref_driver.F90 -- C-callable entry points around the reference's own Fortran P3.
!
! TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
! reference sources where they lie (components/eam/src/physics/p3/scream/
! micro_p3.F90 and its modules) into oracle/_ref/libp3ref.so.  Used in this
! container to pin the C restatement (p3_oracle.c) and to write golden vectors.
!
! Initialisation mirrors the C++ test bridge: micro_p3_utils_init receives the
! C++ physics constants (components/eamxx/src/physics/p3/p3_f90.cpp:96-101),
! then p3_init reads the ice table text and builds the rain tables
! (micro_p3.F90:91-112).
module p3ref_driver
  use iso_c_binding
  use physics_utils, only: rtype
  use micro_p3_utils, only: micro_p3_utils_init
  use micro_p3, only: p3_init, p3_main, p3_get_tables, ice_table_vals, collect_table_vals, &
       find_lookupTable_indices_1a, find_lookupTable_indices_1b, access_lookup_table, &
       access_lookup_table_coll
  implicit none

contains

  ! physics_constants.hpp:27-46 values, passed the way p3_f90.cpp:96-101 does.
  function p3ref_init(dir_c, ndir) result(info) bind(c, name="p3ref_init")
    character(kind=c_char), dimension(*), intent(in) :: dir_c
    integer(c_int), value, intent(in) :: ndir
    integer(c_int) :: info
    character(len=1024) :: dir
    character(len=16) :: version
    integer :: i
    logical :: master
    dir = ' '
    do i = 1, ndir
       dir(i:i) = dir_c(i)
    end do
    master = .false.
    call micro_p3_utils_init(1004.64_rtype, 287.042_rtype, 461.505_rtype, 1000.0_rtype, &
         18.016_rtype, 28.966_rtype, 9.80616_rtype, 2501000.0_rtype, 333700.0_rtype, &
         4188.0_rtype, 273.15_rtype, 3.14159265358979323_rtype, 6, master)
    version = '4.1.1'
    call p3_init(trim(dir), version)
    info = 0
  end function p3ref_init

  ! Tables in C row-major order: ice [5][4][50][12], coll [5][4][50][30][2],
  ! rain [300][10], mu_r [150].
  subroutine p3ref_get_tables(ice, coll, vn, vm, revap, mur) bind(c, name="p3ref_get_tables")
    real(c_double), intent(out) :: ice(12, 50, 4, 5), coll(2, 30, 50, 4, 5)
    real(c_double), intent(out) :: vn(10, 300), vm(10, 300), revap(10, 300), mur(150)
    real(rtype) :: mu_r_u(150), vn_u(300, 10), vm_u(300, 10), revap_u(300, 10)
    integer :: jj, ii, i, j, k
    do jj = 1, 5
       do ii = 1, 4
          do i = 1, 50
             do j = 1, 12
                ice(j, i, ii, jj) = ice_table_vals(jj, ii, i, j)
             end do
             do j = 1, 30
                do k = 1, 2
                   coll(k, j, i, ii, jj) = collect_table_vals(jj, ii, i, j, k)
                end do
             end do
          end do
       end do
    end do
    call p3_get_tables(mu_r_u, revap_u, vn_u, vm_u)
    do i = 1, 300
       do j = 1, 10
          vn(j, i) = vn_u(i, j)
          vm(j, i) = vm_u(i, j)
          revap(j, i) = revap_u(i, j)
       end do
    end do
    mur = mu_r_u
  end subroutine p3ref_get_tables

  ! Ice-table lookup + interpolation (micro_p3.F90:1506-1675), 0-based indices out.
  subroutine p3ref_lookup_ice(qi, ni, qm, rhop, idx, dumi, dumii, dumjj, dum1, dum4, dum5, proc) &
       bind(c, name="p3ref_lookup_ice")
    real(c_double), value, intent(in) :: qi, ni, qm, rhop
    integer(c_int), value, intent(in) :: idx
    integer(c_int), intent(out) :: dumi, dumii, dumjj
    real(c_double), intent(out) :: dum1, dum4, dum5, proc
    integer :: i1, ii1, jj1, zz
    real(rtype) :: d1, d4, d5, d6, p
    call find_lookupTable_indices_1a(i1, jj1, ii1, zz, d1, d4, d5, d6, 50, 4, 5, qi, ni, qm, rhop)
    call access_lookup_table(jj1, ii1, i1, idx + 1, d1, d4, d5, p)
    dumi = i1 - 1; dumii = ii1 - 1; dumjj = jj1 - 1
    dum1 = d1; dum4 = d4; dum5 = d5; proc = p
  end subroutine p3ref_lookup_ice

  ! Ice-rain collection lookup (micro_p3.F90:1545-1606, 1679-1718).
  subroutine p3ref_lookup_coll(qi, ni, qm, rhop, qr, nr, idx, dumj, dum3, proc) &
       bind(c, name="p3ref_lookup_coll")
    real(c_double), value, intent(in) :: qi, ni, qm, rhop, qr, nr
    integer(c_int), value, intent(in) :: idx
    integer(c_int), intent(out) :: dumj
    real(c_double), intent(out) :: dum3, proc
    integer :: i1, ii1, jj1, zz, j1
    real(rtype) :: d1, d3, d4, d5, d6, p
    call find_lookupTable_indices_1a(i1, jj1, ii1, zz, d1, d4, d5, d6, 50, 4, 5, qi, ni, qm, rhop)
    call find_lookupTable_indices_1b(j1, d3, 30, qr, nr)
    call access_lookup_table_coll(jj1, ii1, j1, i1, idx + 1, d1, d3, d4, d5, p)
    dumj = j1 - 1; dum3 = d3; proc = p
  end subroutine p3ref_lookup_coll

  ! One p3_main call (micro_p3.F90:1126-1502).  Arrays are (ncol, nlev)
  ! column-fastest, i.e. level-major; flux arrays have nlev+1 levels.
  subroutine p3ref_main(ncol, nlev, dt, it, predict_nc, prescribed_ccn, &
       qc, nc, qr, nr, th, qv, qi, qm, ni, bm, pres, dz, nc_nuceat_tend, nccn, &
       ni_activated, inv_qc_relvar, precip_liq_surf, precip_ice_surf, &
       diag_eff_radius_qc, diag_eff_radius_qi, diag_eff_radius_qr, rho_qi, dpres, inv_exner, &
       qv2qi_depos_tend, precip_total_tend, nevapr, qr_evap_tend, precip_liq_flux, &
       precip_ice_flux, cld_frac_r, cld_frac_l, cld_frac_i, mu_c, lamc, liq_ice_exchange, &
       vap_liq_exchange, vap_ice_exchange, qv_prev, t_prev) bind(c, name="p3ref_main")
    integer(c_int), value, intent(in) :: ncol, nlev, it, predict_nc, prescribed_ccn
    real(c_double), value, intent(in) :: dt
    real(c_double), dimension(ncol, nlev), intent(inout) :: qc, nc, qr, nr, th, qv, qi, qm, ni, bm
    real(c_double), dimension(ncol, nlev), intent(in) :: pres, dz, nc_nuceat_tend, nccn, ni_activated, &
         inv_qc_relvar, dpres, inv_exner, cld_frac_r, cld_frac_l, cld_frac_i, qv_prev, t_prev
    real(c_double), dimension(ncol), intent(out) :: precip_liq_surf, precip_ice_surf
    real(c_double), dimension(ncol, nlev), intent(out) :: diag_eff_radius_qc, diag_eff_radius_qi, &
         diag_eff_radius_qr, rho_qi, qv2qi_depos_tend, precip_total_tend, nevapr, qr_evap_tend, &
         mu_c, lamc, liq_ice_exchange, vap_liq_exchange, vap_ice_exchange
    real(c_double), dimension(ncol, nlev + 1), intent(out) :: precip_liq_flux, precip_ice_flux
    real(rtype), allocatable :: tend_out(:, :, :)
    real(rtype) :: col_location(ncol, 3)
    logical :: pnc, pccn
    allocate(tend_out(ncol, nlev, 49))
    col_location = 0
    pnc = predict_nc /= 0
    pccn = prescribed_ccn /= 0
    call p3_main(qc, nc, qr, nr, th, qv, dt, qi, qm, ni, bm, &
         pres, dz, nc_nuceat_tend, nccn, ni_activated, inv_qc_relvar, it, precip_liq_surf, &
         precip_ice_surf, 1, ncol, 1, nlev, diag_eff_radius_qc, diag_eff_radius_qi, &
         diag_eff_radius_qr, rho_qi, pnc, pccn, dpres, inv_exner, qv2qi_depos_tend, &
         precip_total_tend, nevapr, qr_evap_tend, precip_liq_flux, precip_ice_flux, &
         cld_frac_r, cld_frac_l, cld_frac_i, tend_out, mu_c, lamc, liq_ice_exchange, &
         vap_liq_exchange, vap_ice_exchange, qv_prev, t_prev, col_location)
    deallocate(tend_out)
  end subroutine p3ref_main

  ! The elementary functions as micro_p3.F90 evaluates them in this build
  ! (bfb_math.inc without SCREAM_CONFIG_IS_CMAKE: the Fortran intrinsics), so
  ! that tests can compare them with the oracle's C library calls.
  real(c_double) function p3ref_gamma(x) bind(c, name="p3ref_gamma")
    real(c_double), value, intent(in) :: x
    p3ref_gamma = gamma(x)
  end function p3ref_gamma

  real(c_double) function p3ref_pow(x, y) bind(c, name="p3ref_pow")
    real(c_double), value, intent(in) :: x, y
    p3ref_pow = x**y
  end function p3ref_pow

end module p3ref_driver
