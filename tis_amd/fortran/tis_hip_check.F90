! tis_hip_check -- a Fortran host program on the C-ABI: reads one DT15
! system (a little-endian stream file: header, parameters, topology arrays,
! coordinates), builds tis_system_c from Fortran arrays, and calls
! neighbor / force_sumup / energy_sumup (and optionally Langevin steps)
! through module tis_hip.  Writes forces (3,N), energies(25) and, after
! nsteps steps, the coordinates to the output file.
!
! usage: tis_hip_check <system file> <output file>
program tis_hip_check
  use, intrinsic :: iso_c_binding
  use tis_hip
  implicit none
  integer(c_int32_t) :: hdr(16)
  real(c_double) :: par(17)
  integer(c_int32_t), allocatable, target :: unit(:), mptype(:), iclass(:), bd_mp(:), ba_mp(:), st_mp(:), &
       tst_mp(:), tst2st(:), tst_excl(:), hb_mp(:), hb_site(:), hb_tert(:), hbval(:)
  real(c_double), allocatable, target :: exv_rad(:), exv_eps(:), charge(:), rl2(:), bd_nat(:), bd_coef(:), &
       ba_nat(:), ba_coef(:), st_nat(:), st_coef(:), tst_nat(:), tst_coef(:), hb_nat(:), hb_coef(:), &
       mass(:), fric(:), radius(:), xyz(:,:), force(:,:), velo(:,:), energy(:), eunit(:,:,:)
  type(tis_system_c), target :: sys(1)
  type(c_ptr) :: ctx
  character(len=4096) :: fin, fout
  integer :: u, n, nu, nbd, nba, nst, ntst, nhb, nsite, nsteps

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) hdr, par
  n = hdr(1); nu = hdr(2); nbd = hdr(3); nba = hdr(4); nst = hdr(5); ntst = hdr(6); nhb = hdr(7)
  nsite = hdr(8); nsteps = hdr(16)
  allocate(unit(n), mptype(n), iclass(n), exv_rad(n), exv_eps(n), charge(n), rl2(nu*nu))
  read(u) unit, mptype, iclass, exv_rad, exv_eps, charge, rl2
  allocate(bd_mp(2*nbd), bd_nat(nbd), bd_coef(nbd), ba_mp(3*nba), ba_nat(nba), ba_coef(nba))
  read(u) bd_mp, bd_nat, bd_coef, ba_mp, ba_nat, ba_coef
  allocate(st_mp(7*nst), st_nat(3*nst), st_coef(4*nst))
  read(u) st_mp, st_nat, st_coef
  allocate(tst_mp(6*ntst), tst_nat(6*ntst), tst_coef(7*ntst), tst2st(2*ntst), tst_excl(2*ntst))
  read(u) tst_mp, tst_nat, tst_coef, tst2st, tst_excl
  allocate(hb_mp(6*nhb), hb_nat(6*nhb), hb_coef(7*nhb), hb_site(6*nhb), hb_tert(nhb), hbval(nsite))
  read(u) hb_mp, hb_nat, hb_coef, hb_site, hb_tert, hbval
  allocate(mass(n), fric(n), radius(n), xyz(3, n))
  read(u) mass, fric, radius, xyz
  close(u)

  sys(1)%nmp = n; sys(1)%nunit = nu
  sys(1)%unit = c_loc(unit); sys(1)%mptype = c_loc(mptype); sys(1)%iclass = c_loc(iclass)
  sys(1)%exv_rad = c_loc(exv_rad); sys(1)%exv_eps = c_loc(exv_eps); sys(1)%charge = c_loc(charge)
  sys(1)%rlist2_unit = c_loc(rl2)
  sys(1)%nbd = nbd; sys(1)%bd_mp = c_loc(bd_mp); sys(1)%bd_nat = c_loc(bd_nat); sys(1)%bd_coef = c_loc(bd_coef)
  sys(1)%nba = nba; sys(1)%ba_mp = c_loc(ba_mp); sys(1)%ba_nat = c_loc(ba_nat); sys(1)%ba_coef = c_loc(ba_coef)
  sys(1)%nst = nst; sys(1)%st_mp = c_loc(st_mp); sys(1)%st_nat = c_loc(st_nat); sys(1)%st_coef = c_loc(st_coef)
  sys(1)%ntst = ntst; sys(1)%tst_mp = c_loc(tst_mp); sys(1)%tst_nat = c_loc(tst_nat)
  sys(1)%tst_coef = c_loc(tst_coef); sys(1)%tst2st = c_loc(tst2st); sys(1)%tst_excl = c_loc(tst_excl)
  sys(1)%nhb = nhb; sys(1)%hb_mp = c_loc(hb_mp); sys(1)%hb_nat = c_loc(hb_nat); sys(1)%hb_coef = c_loc(hb_coef)
  sys(1)%hb_site = c_loc(hb_site); sys(1)%hb_tertiary = c_loc(hb_tert)
  sys(1)%nhbsite = nsite; sys(1)%hbsite_valence = c_loc(hbval)
  sys(1)%mass = c_loc(mass); sys(1)%fric = c_loc(fric); sys(1)%radius = c_loc(radius)
  sys(1)%tempk = par(15)
  sys(1)%p%exv_adjust = par(1); sys(1)%p%exv_dist = par(2); sys(1)%p%exv_dist_PS = par(3)
  sys(1)%p%exv_inf = par(4); sys(1)%p%hb_cutoff_dist = par(5); sys(1)%p%neigh_margin = par(6)
  sys(1)%p%cutoff_ele = par(7); sys(1)%p%ele_coef = par(8); sys(1)%p%ele_cdist = par(9)
  sys(1)%p%ewld_alpha = par(10); sys(1)%p%beta_hb = par(11); sys(1)%p%box = par(12:14)
  sys(1)%p%ele_form = hdr(9); sys(1)%p%dh_cutoff_type = hdr(10)
  sys(1)%p%n_sep_P = hdr(11); sys(1)%p%n_sep_S = hdr(12); sys(1)%p%n_sep_B = hdr(13)
  sys(1)%p%ewld_hmax = hdr(14); sys(1)%p%periodic = hdr(15)

  call hip_check(c_null_ptr, tis_ctx_create(0_c_int32_t, ctx), 'tis_ctx_create')
  call hip_check(ctx, tis_system_upload(ctx, 1_c_int32_t, sys, c_null_ptr, 0_c_int64_t), 'tis_system_upload')
  call hip_check(ctx, tis_state_set(ctx, 0_c_int32_t, 0_c_int32_t, xyz), 'tis_state_set')
  allocate(force(3, n), velo(3, n), energy(TIS_E_MAX), eunit(nu, nu, TIS_E_MAX))
  velo(:,:) = 0.0_c_double
  call hip_neighbor(ctx, 1)
  call hip_force_sumup(ctx, force, 1)
  call hip_energy_sumup(ctx, 1, velo, energy, eunit)
  if (nsteps > 0) then
     call hip_check(ctx, tis_langevin_init(ctx, par(16), 10_c_int32_t, 0_c_int64_t, 1_c_int32_t), 'tis_langevin_init')
     call hip_check(ctx, tis_langevin_steps(ctx, int(nsteps, c_int64_t), TIS_MODE_GRAPH), 'tis_langevin_steps')
     call hip_check(ctx, tis_state_get(ctx, 0_c_int32_t, 0_c_int32_t, xyz), 'tis_state_get')
  end if
  call hip_check(ctx, tis_ctx_destroy(ctx), 'tis_ctx_destroy')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) force, energy
  if (nsteps > 0) write(u) xyz
  close(u)
  write(*, '(a, i0, a, es12.5)') 'tis_hip_check: N=', n, ' E_total=', energy(1)
end program tis_hip_check
