! ref_glue.F90 -- test harness that drives the reference's own Fortran
! routines (compiled from the reference's src/ by oracle/Makefile into
! oracle/_ref/libtisref.so).  TEST INFRASTRUCTURE ONLY: it fills the
! reference's module globals from a tis_system (include/tis_system.h) and
! calls neighbor_list/neighbor_assign/neighbor_list_ele/neighbor_list_hb,
! the force_* and energy_* term routines and simu_ele_set/simu_set_dtrna,
! so that the C restatement in tis_oracle.c can be compared with the
! reference code itself.  One replica (irep = grep = 1), one thread.
module tisref_glue
  use, intrinsic :: iso_c_binding
  use const_maxsize
  use const_index
  use const_physical
  use var_struct
  use var_setp
  use var_simu
  use var_replica
  use var_io
  use mpiconst
  use if_energy
  use mt_stream, only: mt_state, new, init, set_mt19937, genrand_int32, delete
  implicit none

  type, bind(C) :: tis_params_c
     real(c_double) :: exv_adjust, exv_dist, exv_dist_PS, exv_inf
     integer(c_int32_t) :: n_sep_P, n_sep_S, n_sep_B
     real(c_double) :: hb_cutoff_dist, neigh_margin
     integer(c_int32_t) :: ele_form, dh_cutoff_type
     real(c_double) :: cutoff_ele, ele_coef, ele_cdist, ewld_alpha
     integer(c_int32_t) :: ewld_hmax
     real(c_double) :: beta_hb
     integer(c_int32_t) :: periodic
     real(c_double) :: box(3)
  end type tis_params_c

  type, bind(C) :: tis_system_c
     integer(c_int32_t) :: nmp, nunit
     type(c_ptr) :: unit, mptype, iclass, exv_rad, exv_eps, charge, rlist2_unit
     integer(c_int32_t) :: nbd
     type(c_ptr) :: bd_mp, bd_nat, bd_coef
     integer(c_int32_t) :: nba
     type(c_ptr) :: ba_mp, ba_nat, ba_coef
     integer(c_int32_t) :: nst
     type(c_ptr) :: st_mp, st_nat, st_coef
     integer(c_int32_t) :: ntst
     type(c_ptr) :: tst_mp, tst_nat, tst_coef, tst2st, tst_excl
     integer(c_int32_t) :: nhb
     type(c_ptr) :: hb_mp, hb_nat, hb_coef, hb_site, hb_tertiary
     integer(c_int32_t) :: nhbsite
     type(c_ptr) :: hbsite_valence
     type(c_ptr) :: mass, fric, radius
     real(c_double) :: tempk
     type(tis_params_c) :: p
  end type tis_system_c

  integer, save :: g_nexv_cap = 0
  real(c_double), allocatable, save :: g_mass(:)
  integer, save :: g_null = -1
  logical, save :: g_null_open = .false.

contains

  subroutine tisref_load(s) bind(C, name="tisref_load")
    type(tis_system_c), intent(in) :: s
    integer(c_int32_t), pointer :: ia(:), ib(:,:)
    real(c_double), pointer :: da(:), db(:,:)
    integer :: n, i, k, ncharge_tmp

    n = s%nmp
    nmp_all = n; nmp_real = n
    nunit_all = s%nunit; nunit_real = s%nunit
    nthreads = 1; npar_mpi = 1; local_rank_mpi = 0; myrank = 0
    nmp_l = n
    if (allocated(imp_l2g)) deallocate(imp_l2g)
    allocate(imp_l2g(n))
    do i = 1, n
       imp_l2g(i) = i
    end do
    n_replica_all = 1; n_replica_mpi = 1
    irep2grep(1) = 1
    istep = 0
    outfile%data = 6
    if (.not. g_null_open) then
       open(newunit=g_null, file='/dev/null', status='old', action='write')
       g_null_open = .true.
    end if
    outfile%data = g_null

    call c_f_pointer(s%unit, ia, [n]);   imp2unit(1:n) = ia + 1
    call c_f_pointer(s%mptype, ia, [n]); imp2type(1:n) = ia
    call c_f_pointer(s%iclass, ia, [n]); iclass_mp(1:n) = ia
    call c_f_pointer(s%exv_rad, da, [n]); exv_radius_mp(1:n) = da
    call c_f_pointer(s%exv_eps, da, [n]); exv_epsilon_mp(1:n) = da
    ires_mp(1:n) = 0
    do i = 1, s%nunit
       lunit2mp(1, i) = huge(1); lunit2mp(2, i) = 0
    end do
    do i = 1, n
       k = imp2unit(i)
       lunit2mp(1, k) = min(lunit2mp(1, k), i)
       lunit2mp(2, k) = max(lunit2mp(2, k), i)
       iclass_unit(k) = iclass_mp(i)
    end do

    ! parameters
    indtrna%exv_adjust = s%p%exv_adjust
    indtrna%exv_dist = s%p%exv_dist
    indtrna%exv_dist_PS = s%p%exv_dist_PS
    indtrna%exv_inf = s%p%exv_inf
    indtrna%n_sep_nlocal_P = s%p%n_sep_P
    indtrna%n_sep_nlocal_S = s%p%n_sep_S
    indtrna%n_sep_nlocal_B = s%p%n_sep_B
    indtrna%hb_cutoff_dist = s%p%hb_cutoff_dist
    indtrna15 = indtrna
    insimu%neigh_margin = s%p%neigh_margin
    inmisc%i_dtrna_model = 2015
    inmisc%i_temp_independent = 0
    inmisc%i_exv_from_file = 0
    inmisc%i_neigh_dynamic = 1
    inmisc%i_BBR = 0
    inmisc%force_flag(:) = .false.
    inmisc%force_flag(INTERACT%EXV_DT15) = .true.
    inmisc%force_flag(INTERACT%ELE) = (s%p%ele_form >= 0)
    inmisc%class_flag(:) = .false.
    inmisc%class_flag(CLASS%RNA) = .true.
    inmisc%flag_nlocal_unit(:,:,:) = .false.
    call c_f_pointer(s%rlist2_unit, db, [s%nunit, s%nunit])
    do i = 1, s%nunit
       do k = 1, s%nunit
          inmisc%flag_nlocal_unit(i, k, INTERACT%EXV_DT15) = .true.
          inmisc%flag_nlocal_unit(i, k, INTERACT%ELE) = (s%p%ele_form >= 0)
          inmisc%flag_nlocal_unit(i, k, INTERACT%DTRNA) = .true.
          inmisc%rneighbordist2_unit(i, k) = db(i, k)
       end do
    end do
    inpro%n_sep_nlocal = 3
    inperi%i_periodic = s%p%periodic
    inperi%psize(1:3) = s%p%box
    inperi%psizeh(1:3) = inperi%psize(1:3)/2.0
    inperi%d_mirror(:,:) = 0.0e0_PREC
    if (s%p%periodic == 1) call fill_mirror()
    inele%i_function_form = s%p%ele_form
    inele%i_DH_cutoff_type = s%p%dh_cutoff_type
    inele%cutoff_ele = s%p%cutoff_ele
    inele%cdist(1) = s%p%ele_cdist
    inele%coef(1) = s%p%ele_coef
    inele%i_sopsc_consec_res = 1
    inele%n_charge_change = 0
    inele%ewld_dipole = -1
    beta_hbond15 = s%p%beta_hb

    ! charges: every bead with a charge value, in bead order
    call c_f_pointer(s%charge, da, [n])
    ncharge_tmp = 0
    do i = 1, n
       if (da(i) /= 0.0e0_PREC) ncharge_tmp = ncharge_tmp + 1
    end do
    ncharge = ncharge_tmp
    if (allocated(icharge2mp)) deallocate(icharge2mp, lmp2charge, coef_charge)
    allocate(icharge2mp(max(ncharge,1)), lmp2charge(n), coef_charge(max(ncharge,1), 1))
    lmp2charge(:) = 0
    k = 0
    do i = 1, n
       if (da(i) /= 0.0e0_PREC) then
          k = k + 1
          icharge2mp(k) = i
          lmp2charge(i) = k
          coef_charge(k, 1) = da(i)
       end if
    end do

    ! bonds and angles
    if (allocated(ibd2mp)) deallocate(ibd2mp, bd_nat, coef_bd)
    nbd = s%nbd
    allocate(ibd2mp(2, max(nbd,1)), bd_nat(max(nbd,1)), coef_bd(max(nbd,1)))
    if (nbd > 0) then
       call c_f_pointer(s%bd_mp, ib, [2, nbd]); ibd2mp(:, 1:nbd) = ib + 1
       call c_f_pointer(s%bd_nat, da, [nbd]); bd_nat(1:nbd) = da
       call c_f_pointer(s%bd_coef, da, [nbd]); coef_bd(1:nbd) = da
    end if
    if (allocated(iba2mp)) deallocate(iba2mp, ba_nat, coef_ba)
    nba = s%nba
    allocate(iba2mp(3, max(nba,1)), ba_nat(max(nba,1)), coef_ba(max(nba,1)))
    if (nba > 0) then
       call c_f_pointer(s%ba_mp, ib, [3, nba]); iba2mp(:, 1:nba) = ib + 1
       call c_f_pointer(s%ba_nat, da, [nba]); ba_nat(1:nba) = da
       call c_f_pointer(s%ba_coef, da, [nba]); coef_ba(1:nba) = da
    end if

    ! stacks
    if (allocated(idtrna_st2mp)) deallocate(idtrna_st2mp, dtrna_st_nat, coef_dtrna_st)
    ndtrna_st = s%nst
    allocate(idtrna_st2mp(7, max(ndtrna_st,1)), dtrna_st_nat(3, max(ndtrna_st,1)), &
             coef_dtrna_st(0:3, max(ndtrna_st,1), 1))
    if (ndtrna_st > 0) then
       call c_f_pointer(s%st_mp, ib, [7, ndtrna_st]); idtrna_st2mp(:, 1:ndtrna_st) = ib + 1
       call c_f_pointer(s%st_nat, db, [3, ndtrna_st]); dtrna_st_nat(:, 1:ndtrna_st) = db
       call c_f_pointer(s%st_coef, db, [4, ndtrna_st]); coef_dtrna_st(0:3, 1:ndtrna_st, 1) = db
    end if
    if (allocated(st_status)) deallocate(st_status)
    allocate(st_status(max(ndtrna_st,1), 1))
    st_status(:,:) = .true.

    ! tertiary stacks
    if (allocated(idtrna_tst2mp)) deallocate(idtrna_tst2mp, dtrna_tst_nat, coef_dtrna_tst, &
                                             idtrna_tst2st, flg_tst_exclusive)
    ndtrna_tst = s%ntst
    allocate(idtrna_tst2mp(6, max(ndtrna_tst,1)), dtrna_tst_nat(6, max(ndtrna_tst,1)), &
             coef_dtrna_tst(0:6, max(ndtrna_tst,1)), idtrna_tst2st(2, max(ndtrna_tst,1)), &
             flg_tst_exclusive(2, max(ndtrna_tst,1)))
    if (ndtrna_tst > 0) then
       call c_f_pointer(s%tst_mp, ib, [6, ndtrna_tst]); idtrna_tst2mp(:, 1:ndtrna_tst) = ib + 1
       call c_f_pointer(s%tst_nat, db, [6, ndtrna_tst]); dtrna_tst_nat(:, 1:ndtrna_tst) = db
       call c_f_pointer(s%tst_coef, db, [7, ndtrna_tst]); coef_dtrna_tst(0:6, 1:ndtrna_tst) = db
       call c_f_pointer(s%tst2st, ib, [2, ndtrna_tst]); idtrna_tst2st(:, 1:ndtrna_tst) = ib + 1
       call c_f_pointer(s%tst_excl, ib, [2, ndtrna_tst]); flg_tst_exclusive(:, 1:ndtrna_tst) = (ib /= 0)
    end if

    ! hydrogen bonds
    if (allocated(idtrna_hb2mp)) deallocate(idtrna_hb2mp, dtrna_hb_nat, coef_dtrna_hb, &
         idtrna_hb2hbsite, flg_hb_tertiary, dtrna_hb_neigh_dist2)
    ndtrna_hb = s%nhb
    allocate(idtrna_hb2mp(6, max(ndtrna_hb,1)), dtrna_hb_nat(6, max(ndtrna_hb,1)), &
             coef_dtrna_hb(0:6, max(ndtrna_hb,1)), idtrna_hb2hbsite(3, 2, max(ndtrna_hb,1)), &
             flg_hb_tertiary(max(ndtrna_hb,1)), dtrna_hb_neigh_dist2(max(ndtrna_hb,1)))
    idtrna_hb2hbsite(:,:,:) = 0
    if (ndtrna_hb > 0) then
       call c_f_pointer(s%hb_mp, ib, [6, ndtrna_hb]); idtrna_hb2mp(:, 1:ndtrna_hb) = ib + 1
       call c_f_pointer(s%hb_nat, db, [6, ndtrna_hb]); dtrna_hb_nat(:, 1:ndtrna_hb) = db
       call c_f_pointer(s%hb_coef, db, [7, ndtrna_hb]); coef_dtrna_hb(0:6, 1:ndtrna_hb) = db
       call c_f_pointer(s%hb_site, ib, [6, ndtrna_hb])
       do i = 1, ndtrna_hb
          do k = 1, 3
             idtrna_hb2hbsite(k, 1, i) = ib(2*k-1, i) + 1
             idtrna_hb2hbsite(k, 2, i) = ib(2*k, i) + 1
          end do
       end do
       call c_f_pointer(s%hb_tertiary, ia, [ndtrna_hb]); flg_hb_tertiary(1:ndtrna_hb) = (ia /= 0)
       ! mloop_dtrna.F90:149-151 (neighbour-dynamic radius)
       do i = 1, ndtrna_hb
          dtrna_hb_neigh_dist2(i) = (dtrna_hb_nat(1,i) + indtrna%hb_cutoff_dist + insimu%neigh_margin) ** 2
       end do
    end if
    nhbsite = s%nhbsite
    if (allocated(nvalence_hbsite)) deallocate(nvalence_hbsite, list_hb_at_hbsite, num_hb_at_hbsite)
    allocate(nvalence_hbsite(max(nhbsite,1)), list_hb_at_hbsite(MXMPHBNEIGHBOR, max(nhbsite,1), 1), &
             num_hb_at_hbsite(max(nhbsite,1), 1))
    if (nhbsite > 0) then
       call c_f_pointer(s%hbsite_valence, ia, [nhbsite]); nvalence_hbsite(1:nhbsite) = ia
    end if
    if (allocated(hb_status)) deallocate(hb_status, hbsite_excess, for_hb, ene_hb)
    allocate(hb_status(max(ndtrna_hb,1), 1), hbsite_excess(max(nhbsite,1)), &
             for_hb(3, 6, max(ndtrna_hb,1)), ene_hb(max(ndtrna_hb,1), 1))
    hb_status(:,:) = .false.
    if (allocated(nhbneigh)) deallocate(nhbneigh, ineigh2hb)
    allocate(nhbneigh(1), ineigh2hb(max(MXMPHBNEIGHBOR*n/2, ndtrna_hb, 1), 1))
    nhbneigh(:) = 0

    ! coordinates and lists
    if (allocated(xyz_mp_rep)) deallocate(xyz_mp_rep, pxyz_mp_rep)
    allocate(xyz_mp_rep(3, n, 1), pxyz_mp_rep(3, n, 1))
    g_nexv_cap = MXMPNEIGHBOR * n
    if (allocated(lexv)) deallocate(lexv, iexv2mp)
    allocate(lexv(2, E_TYPE%MAX, 1), iexv2mp(3, g_nexv_cap, 1))
    if (allocated(lele)) deallocate(lele, iele2mp, coef_ele)
    allocate(lele(1), iele2mp(4, max(MXMPELE*ncharge,1), 1), coef_ele(max(MXMPELE*ncharge,1), 1))
    lele(:) = 0
    ! Ewald tables (mloop_ewld.F90), cubic box
    if (s%p%ele_form == 2) then
       inele%ewld_alpha = s%p%ewld_alpha
       inele%ewld_hmax = real(s%p%ewld_hmax, PREC)
       if (allocated(ewld_h)) deallocate(ewld_h)
       if (allocated(ewld_f_coef)) deallocate(ewld_f_coef)
       if (allocated(ewld_f_rlv)) deallocate(ewld_f_rlv)
       call mloop_ewld()
    end if
    if (allocated(lmp2con)) deallocate(lmp2con, lmp2LJ, lmp2wca)
    allocate(lmp2con(n), lmp2LJ(n), lmp2wca(n))
    lmp2con(:) = 0; lmp2LJ(:) = 0; lmp2wca(:) = 0
  end subroutine tisref_load

  subroutine fill_mirror()
    integer :: ix, iy, iz, im
    im = 0
    do ix = 1, 3
       do iy = 1, 3
          do iz = 1, 3
             im = im + 1
             inperi%d_mirror(1, im) = merge(-inperi%psize(1), merge(inperi%psize(1), 0.0e0_PREC, ix == 2), ix == 1)
             inperi%d_mirror(2, im) = merge(-inperi%psize(2), merge(inperi%psize(2), 0.0e0_PREC, iy == 2), iy == 1)
             inperi%d_mirror(3, im) = merge(-inperi%psize(3), merge(inperi%psize(3), 0.0e0_PREC, iz == 2), iz == 1)
          end do
       end do
    end do
  end subroutine fill_mirror

  ! Set coordinates (3,N) and rebuild every list as neighbor.F90:35-76 does.
  subroutine tisref_neighbor(xyz) bind(C, name="tisref_neighbor")
    real(c_double), intent(in) :: xyz(3, nmp_all)
    integer :: lmp2neigh((nmp_l+nthreads-1)/nthreads, 0:nthreads-1)
    integer, allocatable :: ineigh2mp(:,:)
    xyz_mp_rep(:,:,1) = xyz
    pxyz_mp_rep(:,:,1) = xyz
    if (inperi%i_periodic == 1) call util_periodic(1)
    allocate(ineigh2mp(MXMPNEIGHBOR*nmp_all/nthreads, 0:nthreads-1))
    call neighbor_list(1, ineigh2mp, lmp2neigh)
    call neighbor_assign(1, ineigh2mp, lmp2neigh)
    deallocate(ineigh2mp)
    if (inmisc%force_flag(INTERACT%ELE)) call neighbor_list_ele(1)
    call neighbor_list_hb(1)
  end subroutine tisref_neighbor

  ! Move the coordinates without rebuilding the lists (pxyz follows xyz by the
  ! same displacement, as in time_integral.F90:156-160).
  subroutine tisref_move(xyz) bind(C, name="tisref_move")
    real(c_double), intent(in) :: xyz(3, nmp_all)
    pxyz_mp_rep(:,:,1) = pxyz_mp_rep(:,:,1) + (xyz - xyz_mp_rep(:,:,1))
    xyz_mp_rep(:,:,1) = xyz
  end subroutine tisref_move

  function tisref_list_count(which) result(n) bind(C, name="tisref_list_count")
    integer(c_int32_t), value :: which
    integer(c_int64_t) :: n
    select case (which)
    case (0)
       n = lexv(2, E_TYPE%EXV_DT15, 1) - lexv(1, E_TYPE%EXV_DT15, 1) + 1
    case (1)
       n = lele(1)
    case default
       n = nhbneigh(1)
    end select
  end function tisref_list_count

  ! Copy a list out: which 0 = exv (i, j, mirror) 0-based, 1 = ele (+coef),
  ! 2 = hb candidate indices.
  subroutine tisref_list_get(which, pairs, coef) bind(C, name="tisref_list_get")
    integer(c_int32_t), value :: which
    integer(c_int32_t), intent(out) :: pairs(*)
    real(c_double), intent(out) :: coef(*)
    integer :: k, m, ksta
    select case (which)
    case (0)
       ksta = lexv(1, E_TYPE%EXV_DT15, 1)
       m = 0
       do k = ksta, lexv(2, E_TYPE%EXV_DT15, 1)
          pairs(3*m+1) = iexv2mp(1, k, 1) - 1
          pairs(3*m+2) = iexv2mp(2, k, 1) - 1
          pairs(3*m+3) = iexv2mp(3, k, 1) - 1
          m = m + 1
       end do
    case (1)
       do k = 1, lele(1)
          pairs(3*k-2) = iele2mp(1, k, 1) - 1
          pairs(3*k-1) = iele2mp(2, k, 1) - 1
          pairs(3*k)   = iele2mp(3, k, 1) - 1
          coef(k) = coef_ele(k, 1)
       end do
    case default
       do k = 1, nhbneigh(1)
          pairs(k) = ineigh2hb(k, 1) - 1
       end do
    end select
  end subroutine tisref_list_get

  ! One force term into f(3,N) (zeroed first).  term: 0 bond, 1 angle,
  ! 2 tstack+stack (st_status from the tstack pass), 3 stack only (all
  ! st_status true), 4 hbond15, 5 exv, 6 DH, 7 plain Coulomb, 8 Ewald.
  subroutine tisref_force_term(term, f) bind(C, name="tisref_force_term")
    integer(c_int32_t), value :: term
    real(c_double), intent(out) :: f(3, nmp_all)
    logical :: st_l(max(ndtrna_st,1))
    f(:,:) = 0.0e0_PREC
    select case (term)
    case (0)
       call force_bond(1, f)
    case (1)
       call force_bangle(1, f)
    case (2)
       call force_dtrna_stack_nlocal(1, f, st_l)
       st_status(1:ndtrna_st, 1) = st_l(1:ndtrna_st)
       call force_dtrna_stack(1, f)
    case (3)
       st_status(:, 1) = .true.
       call force_dtrna_stack(1, f)
    case (4)
       call force_dtrna_hbond15(1, f)
    case (5)
       call force_exv_dt15(1, f)
    case (6)
       call force_ele_DH(1, f)
    case (7)
       call force_ele_coulomb(1, f)
    case (8)
       call force_ele_coulomb_ewld(1, f)
    end select
  end subroutine tisref_force_term

  ! Diffusion tensor and its Cholesky factor by simu_hydro_tensors for n
  ! beads (n <= MXMP): D and L (3n,3n) column-major.
  subroutine tisref_hydro(n, xyz, rad, tempk, visc, htype, D, L) bind(C, name="tisref_hydro")
    integer(c_int32_t), value :: n, htype
    real(c_double), intent(in) :: xyz(3, n), rad(n)
    real(c_double), value :: tempk, visc
    real(c_double), intent(out) :: D(3*n, 3*n), L(3*n, 3*n)
    nmp_real = n
    if (allocated(xyz_mp_rep)) deallocate(xyz_mp_rep, pxyz_mp_rep)
    allocate(xyz_mp_rep(3, n, 1), pxyz_mp_rep(3, n, 1))
    xyz_mp_rep(:,:,1) = xyz
    radius(1:n) = rad
    inpara%viscosity = visc
    inmisc%i_hydro_tensor = htype
    if (allocated(diffuse_tensor)) deallocate(diffuse_tensor)
    if (allocated(random_tensor)) deallocate(random_tensor)
    allocate(diffuse_tensor(3*n, 3*n), random_tensor(3*n, 3*n))
    call simu_hydro_tensors(1, tempk)
    D = diffuse_tensor
    L = random_tensor
  end subroutine tisref_hydro

  ! FENE + WCA forces (force_fene, force_wca) of a non-periodic bead set.
  subroutine tisref_bd_forces(n, xyz, nf, fmp, fnat, fcoef, fdist2, nw, wmp, wnat2, wcoef, wcut, f) &
       bind(C, name="tisref_bd_forces")
    integer(c_int32_t), value :: n, nf, nw
    real(c_double), intent(in) :: xyz(3, n)
    integer(c_int32_t), intent(in) :: fmp(2, *), wmp(2, *)
    real(c_double), intent(in) :: fnat(*), fcoef(*), fdist2(*), wnat2(*), wcoef(2, *)
    real(c_double), value :: wcut
    real(c_double), intent(out) :: f(3, n)
    real(c_double), allocatable :: fl(:,:)
    integer :: k
    nmp_all = n
    if (allocated(xyz_mp_rep)) deallocate(xyz_mp_rep, pxyz_mp_rep)
    allocate(xyz_mp_rep(3, n, 1), pxyz_mp_rep(3, n, 1))
    xyz_mp_rep(:,:,1) = xyz
    pxyz_mp_rep(:,:,1) = xyz
    inperi%i_periodic = 0
    nfene = nf
    if (allocated(ifene2mp)) deallocate(ifene2mp, fene_nat, coef_fene, dist2_fene)
    allocate(ifene2mp(2, max(nf,1)), fene_nat(max(nf,1)), coef_fene(max(nf,1)), dist2_fene(max(nf,1)))
    do k = 1, nf
       ifene2mp(:, k) = fmp(:, k) + 1
       fene_nat(k) = fnat(k); coef_fene(k) = fcoef(k); dist2_fene(k) = fdist2(k)
    end do
    nwca = nw
    if (allocated(iwca2mp)) deallocate(iwca2mp, wca_nat2, coef_wca)
    allocate(iwca2mp(2, max(nw,1)), wca_nat2(max(nw,1)), coef_wca(max(nw,1), 2))
    do k = 1, nw
       iwca2mp(:, k) = wmp(:, k) + 1
       wca_nat2(k) = wnat2(k); coef_wca(k, 1) = wcoef(1, k); coef_wca(k, 2) = wcoef(2, k)
    end do
    inpro%cutoff_wca = wcut
    allocate(fl(3, n))
    fl(:,:) = 0.0e0_PREC
    call force_fene(1, fl)
    f = fl
    fl(:,:) = 0.0e0_PREC
    call force_wca(1, fl)
    f = f + fl
    deallocate(fl)
  end subroutine tisref_bd_forces

  ! Seed the replica's stream mts(1,0), which force_dtrna_hbond15 draws its
  ! valence-pass numbers from (force_dtrna_hbond15.F90:392,413,426), by the
  ! scalar initialisation of mt_stream (as setp_random.F90:29-31 prepares
  ! mts(0,0) before deriving the replica streams).
  subroutine tisref_mt_seed(seed) bind(C, name="tisref_mt_seed")
    integer(c_int32_t), value :: seed
    if (allocated(mts)) deallocate(mts)
    allocate(mts(0:1, 0:0))
    call set_mt19937
    call new(mts(1, 0))
    call init(mts(1, 0), seed)
  end subroutine tisref_mt_seed

  ! The first n 32-bit outputs of a stream seeded as in tisref_mt_seed: the
  ! words the valence pass consumes, two per genrand_double1/_double4 draw.
  subroutine tisref_mt_raw(seed, n, raw) bind(C, name="tisref_mt_raw")
    integer(c_int32_t), value :: seed, n
    integer(c_int32_t), intent(out) :: raw(n)
    type(mt_state) :: st
    integer :: k
    call set_mt19937
    call new(st)
    call init(st, seed)
    do k = 1, n
       raw(k) = genrand_int32(st)
    end do
    call delete(st)
  end subroutine tisref_mt_raw

  subroutine tisref_hb_state(status, ene) bind(C, name="tisref_hb_state")
    integer(c_int32_t), intent(out) :: status(*)
    real(c_double), intent(out) :: ene(*)
    integer :: k
    do k = 1, ndtrna_hb
       status(k) = merge(1, 0, hb_status(k, 1))
       ene(k) = ene_hb(k, 1)
    end do
  end subroutine tisref_hb_state

  ! energies per E_TYPE (0-based slots, TIS_E_MAX entries) for the current
  ! coordinates; velocities optional (nvelo = 0 -> no kinetic term).
  subroutine tisref_energy(velo, nvelo, energy) bind(C, name="tisref_energy")
    real(c_double), intent(in) :: velo(3, *)
    integer(c_int32_t), value :: nvelo
    real(c_double), intent(out) :: energy(E_TYPE%MAX)
    real(PREC) :: e(E_TYPE%MAX), eu(nunit_all, nunit_all, E_TYPE%MAX)
    real(PREC) :: ene_st(max(ndtrna_st,1)), ene_tst(max(ndtrna_tst,1))
    logical :: st_l(max(ndtrna_st,1))
    integer :: i
    e(:) = 0.0e0_PREC
    eu(:,:,:) = 0.0e0_PREC
    call energy_bond(1, eu, e)
    call energy_bangle(1, eu, e)
    call energy_dtrna_stack_nlocal(1, eu, e, ene_tst, st_l)
    st_status(1:ndtrna_st, 1) = st_l(1:ndtrna_st)
    call energy_dtrna_stack(1, eu, e, ene_st)
    call energy_dtrna_hbond15(1, eu, e)
    call energy_exv_dt15(1, eu, e)
    if (inmisc%force_flag(INTERACT%ELE) .and. inele%i_function_form == 0) then
       call energy_ele_DH(1, e, eu)
    else if (inmisc%force_flag(INTERACT%ELE) .and. inele%i_function_form == 1) then
       call energy_ele_coulomb(1, e, eu)
    else if (inmisc%force_flag(INTERACT%ELE) .and. inele%i_function_form == 2) then
       call energy_ele_coulomb_ewld(1, e, eu)
    end if
    if (nvelo > 0) then
       do i = 1, nmp_all
          e(E_TYPE%VELO) = e(E_TYPE%VELO) + 0.5e0_PREC * g_mass(i) * dot_product(velo(:,i), velo(:,i))
       end do
    end if
    energy(:) = 0.0e0_PREC
    do i = 1, E_TYPE%MAX
       energy(i) = e(i)
    end do
    energy(1) = 0.0e0_PREC
    do i = 1, E_TYPE%MAX
       if (i == E_TYPE%TOTAL .or. i == E_TYPE%VELO) cycle
       energy(1) = energy(1) + e(i)
    end do
  end subroutine tisref_energy

  ! energy_unit of the reference's own term routines at the current
  ! coordinates (the kinetic term as energy_kinetic.F90:49-56 adds it), then
  ! the per-unit post-processing of energy_sumup.F90:480-508 for
  ! i_output_energy_style = style.
  subroutine tisref_energy_unit(velo, nvelo, style, eu_out) bind(C, name="tisref_energy_unit")
    real(c_double), intent(in) :: velo(3, *)
    integer(c_int32_t), value :: nvelo, style
    real(c_double), intent(out) :: eu_out(nunit_all, nunit_all, E_TYPE%MAX)
    real(PREC) :: e(E_TYPE%MAX), eu(nunit_all, nunit_all, E_TYPE%MAX), sume, ev
    real(PREC) :: ene_st(max(ndtrna_st,1)), ene_tst(max(ndtrna_tst,1))
    logical :: st_l(max(ndtrna_st,1))
    integer :: i, iunit, junit
    e(:) = 0.0e0_PREC
    eu(:,:,:) = 0.0e0_PREC
    call energy_bond(1, eu, e)
    call energy_bangle(1, eu, e)
    call energy_dtrna_stack_nlocal(1, eu, e, ene_tst, st_l)
    st_status(1:ndtrna_st, 1) = st_l(1:ndtrna_st)
    call energy_dtrna_stack(1, eu, e, ene_st)
    call energy_dtrna_hbond15(1, eu, e)
    call energy_exv_dt15(1, eu, e)
    if (inmisc%force_flag(INTERACT%ELE) .and. inele%i_function_form == 0) then
       call energy_ele_DH(1, e, eu)
    else if (inmisc%force_flag(INTERACT%ELE) .and. inele%i_function_form == 1) then
       call energy_ele_coulomb(1, e, eu)
    else if (inmisc%force_flag(INTERACT%ELE) .and. inele%i_function_form == 2) then
       call energy_ele_coulomb_ewld(1, e, eu)
    end if
    if (nvelo > 0) then
       do i = 1, nmp_all
          ev = 0.5e0_PREC * g_mass(i) * dot_product(velo(:,i), velo(:,i))
          iunit = imp2unit(i)
          eu(iunit, iunit, E_TYPE%VELO) = eu(iunit, iunit, E_TYPE%VELO) + ev
       end do
    end if
    if (style == 0) then
       do iunit = 1, nunit_all
          do junit = 1, iunit - 1
             eu(iunit, iunit, 1:E_TYPE%MAX) = eu(iunit, iunit, 1:E_TYPE%MAX) + 0.5e0_PREC * eu(junit, iunit, 1:E_TYPE%MAX)
          end do
          do junit = iunit + 1, nunit_all
             eu(iunit, iunit, 1:E_TYPE%MAX) = eu(iunit, iunit, 1:E_TYPE%MAX) + 0.5e0_PREC * eu(iunit, junit, 1:E_TYPE%MAX)
          end do
       end do
    else
       do iunit = 1, nunit_all
          do junit = iunit + 1, nunit_all
             do i = 1, E_TYPE%MAX
                sume = eu(iunit, junit, i) + eu(junit, iunit, i)
                eu(iunit, junit, i) = sume
                eu(junit, iunit, i) = sume
             end do
          end do
       end do
    end if
    do i = 1, E_TYPE%MAX
       if (i == E_TYPE%TOTAL .or. i == E_TYPE%VELO) cycle
       eu(1:nunit_all, 1:nunit_all, E_TYPE%TOTAL) = eu(1:nunit_all, 1:nunit_all, E_TYPE%TOTAL) &
            + eu(1:nunit_all, 1:nunit_all, i)
    end do
    eu_out(:,:,:) = eu(:,:,:)
  end subroutine tisref_energy_unit

  ! mass for the kinetic term (the glue keeps its own copy)
  subroutine tisref_set_mass(m) bind(C, name="tisref_set_mass")
    real(c_double), intent(in) :: m(nmp_all)
    if (allocated(g_mass)) deallocate(g_mass)
    allocate(g_mass(nmp_all))
    g_mass = m
  end subroutine tisref_set_mass

  ! simu_ele_set + simu_set_dtrna for one temperature / ionic strength:
  ! returns inele%coef, inele%cdist, the phosphate charge and U0 of every
  ! stack (from h, s, Tm) and beta_hbond15.
  subroutine tisref_temperature(tempk, ionic, hsTm, coef, cdist, zp, u0, beta) &
       bind(C, name="tisref_temperature")
    real(c_double), value :: tempk, ionic
    real(c_double), intent(in) :: hsTm(3, *)
    real(c_double), intent(out) :: coef, cdist, zp, u0(*), beta
    integer :: k
    inele%i_diele = 1
    inele%i_charge = 1
    inele%ewld_dipole = 0
    inele%length_per_unit(:) = 0.0e0_PREC
    inele%length_per_unit(ifunc_seq2id_local()) = 4.38178046e0_PREC
    if (allocated(dtrna_st_hsTm)) deallocate(dtrna_st_hsTm)
    allocate(dtrna_st_hsTm(3, max(ndtrna_st,1)))
    do k = 1, ndtrna_st
       dtrna_st_hsTm(:, k) = hsTm(:, k)
    end do
    do k = 1, ncharge
       coef_charge(k, 1) = 0.0e0_PREC
    end do
    call simu_ele_set(1, tempk, ionic)
    call simu_set_dtrna(1, tempk)
    coef = inele%coef(1)
    cdist = inele%cdist(1)
    zp = 0.0e0_PREC
    do k = 1, ncharge
       if (imp2type(icharge2mp(k)) == MPTYPE%RNA_PHOS) zp = -coef_charge(k, 1)
    end do
    do k = 1, ndtrna_st
       u0(k) = coef_dtrna_st(0, k, 1)
    end do
    beta = beta_hbond15
  end subroutine tisref_temperature

  ! Flush the standard output unit (the reference's warnings, e.g. the force
  ! clamps of force_exv_dt15, are written there).
  subroutine tisref_flush() bind(C, name="tisref_flush")
    flush(6)
  end subroutine tisref_flush

  integer function ifunc_seq2id_local()
    integer :: ifunc_seq2id
    ifunc_seq2id_local = ifunc_seq2id('P  ')
  end function ifunc_seq2id_local

end module tisref_glue
