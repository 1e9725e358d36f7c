!> Fortran driver of the GPU path, modelled on the reference's own test
!> program test/src/Spin1d_SU2/spinSU2_1d.f90 (spin-1/2 Heisenberg, iDMRG
!> LDMRG=40, MDMRG=20, LANC_NEIGEN=2, LANC_DIM_THRESHOLD=128) and its golden
!> energies (tests/golden/spin12_L40_M20_energy.check, copied data).
!> Part 2 binds spMatVec_gpu_main to a procedure pointer of the reference's
!> abstract interface sparse_HxV and checks it against the direct-product
!> loops of spMatVec_direct_main written out here on the host.  Part 3 runs
!> PBCdmrg=T on the spin-1/2 Heisenberg rings of 4 and 6 sites.
program spin_chain_golden
  use iso_c_binding
  use nss_dmrg
  implicit none

  abstract interface
     subroutine sparse_HxV(Nloc, v, Hv)   ! src/DMRG_GLOBAL.f90:41-53
       integer :: Nloc
       real(8), dimension(Nloc) :: v
       real(8), dimension(Nloc) :: Hv
     end subroutine sparse_HxV
  end interface
  procedure(sparse_HxV), pointer :: spHtimesV_p => null()

  type(nss_model) :: model
  type(nss_params) :: par
  type(nss_step_info) :: info
  real(8) :: Hij(2, 2), ref(3), err, maxerr, szerr, szref, sz(2, 2, 1)
  real(8), allocatable :: avo(:, :)
  integer :: i, u, ios, nbad, nchecked
  character(len=512) :: golden, golden_sz

  ! ---------------------------------------------------------------- part 1
  call get_command_argument(1, golden)
  call get_command_argument(2, golden_sz)
  Hij = 0d0
  Hij(1, 1) = 1d0      ! Jp
  Hij(2, 2) = 0.5d0    ! Jx/2
  model%model = NSS_MODEL_SPIN12
  par%ldmrg = 40
  par%mdmrg = 20
  par%neigen = 2
  par%lanc_dim_threshold = 128
  par%lanc_tol = 1d-12
  call init_dmrg_gpu(Hij, model, par)
  call run_dmrg_gpu('i')
  open(newunit=u, file=trim(golden), status="old", action="read")
  maxerr = 0d0
  nbad = 0
  nchecked = 0
  do i = 1, dmrg_gpu_nsteps()
     read(u, *, iostat=ios) ref
     if (ios /= 0) exit
     info = dmrg_gpu_step(i)
     if (info%left_len /= nint(ref(1))) nbad = nbad + 1
     err = max(abs(info%energies(1) / info%L - ref(2)), abs(info%energies(2) / info%L - ref(3)))
     maxerr = max(maxerr, err)
     nchecked = nchecked + 1
  enddo
  close(u)
  write(*, "(A,I3,A,ES10.3)") "golden iDMRG steps:", nchecked, "  max |dE/L| =", maxerr
  if (maxerr > 1d-10 .or. nbad > 0 .or. nchecked /= 39) then
     write(*, "(A)") "FAIL golden energies"
     stop 1
  endif
  ! Measure_DMRG(S_z, pos=1..40) against the reference's sz.check, to the
  ! reference test's own tolerance (test/src/Spin1d_SU2/spinSU2_1d.f90:60)
  if (len_trim(golden_sz) > 0) then
     sz = 0d0
     sz(1, 1, 1) = 0.5d0
     sz(2, 2, 1) = -0.5d0
     call dmrg_gpu_measure(sz, [(i, i = 1, 40)], avo)
     open(newunit=u, file=trim(golden_sz), status="old", action="read")
     szerr = 0d0
     do i = 1, 40
        read(u, *) szref
        szerr = max(szerr, abs(avo(1, i) - szref))
     enddo
     close(u)
     write(*, "(A,ES10.3)") "golden S_z profile: max |d<Sz>| =", szerr
     if (szerr > 1d-8) then
        write(*, "(A)") "FAIL golden S_z"
        stop 1
     endif
  endif
  call finalize_dmrg_gpu()

  ! ---------------------------------------------------------------- part 2
  call check_drop_in()

  ! ---------------------------------------------------------------- part 3
  call check_pbc_ring()
  call nss_gpu_finalize()
  write(*, "(A)") "PASS"

contains

  !> PBCdmrg=T, LDMRG=3, untruncated: a link-'n' step (ring of 4 sites)
  !> then a link-'p' step (ring of 6) against the exact ring energies -2 and
  !> -2.8027756377319946.
  subroutine check_pbc_ring()
    type(nss_model) :: m
    type(nss_params) :: p
    type(nss_step_info) :: s1, s2
    m%model = NSS_MODEL_SPIN12
    p%ldmrg = 3
    p%mdmrg = 200
    p%neigen = 1
    p%lanc_dim_threshold = 8
    p%lanc_tol = 1d-12
    p%pbc = 1
    call init_dmrg_gpu(Hij, m, p)
    call run_dmrg_gpu('i')
    s1 = dmrg_gpu_step(1)
    s2 = dmrg_gpu_step(2)
    write(*, "(A,2F20.15)") "PBC rings E(4), E(6):", s1%energies(1), s2%energies(1)
    if (dmrg_gpu_nsteps() /= 2 .or. s1%L /= 4 .or. s2%L /= 6 .or. s1%link /= ichar('n') .or. &
        s2%link /= ichar('p') .or. abs(s1%energies(1) + 2d0) > 1d-11 .or. &
        abs(s2%energies(1) + 2.8027756377319946d0) > 1d-11) then
       write(*, "(A)") "FAIL PBC rings"
       stop 1
    endif
    call finalize_dmrg_gpu()
  end subroutine check_pbc_ring

  subroutine check_drop_in()
    integer, parameter :: nq = 3, nt = 4
    integer :: Dls(nq), Drs(nq), dst(nt), src(nt), offs(nq)
    type(dense_block) :: HL(nq), HR(nq), A(nt), B(nt)
    real(8), allocatable :: v(:), hv(:), ref(:), C(:,:), evecs(:,:)
    real(8) :: evals(1), res
    integer :: q, t, n, il, ir, jl, jr, ai, aj, bi, bj, nop
    integer, allocatable :: seed(:)
    call random_seed(size=n)
    allocate(seed(n))
    seed = 20260
    call random_seed(put=seed)
    Dls = [3, 70, 5]
    Drs = [65, 4, 33]
    dst = [1, 2, 2, 3]
    src = [2, 1, 3, 3]
    n = 0
    do q = 1, nq
       offs(q) = n
       n = n + Dls(q) * Drs(q)
       allocate(HL(q)%m(Dls(q), Dls(q)), HR(q)%m(Drs(q), Drs(q)))
       call random_number(HL(q)%m)
       call random_number(HR(q)%m)
       HL(q)%m = HL(q)%m + transpose(HL(q)%m)
       HR(q)%m = HR(q)%m + transpose(HR(q)%m)
    enddo
    do t = 1, nt
       allocate(A(t)%m(Dls(dst(t)), Dls(src(t))), B(t)%m(Drs(dst(t)), Drs(src(t))))
       call random_number(A(t)%m)
       call random_number(B(t)%m)
    enddo
    call nss_gpu_set_superblock(Dls, Drs, HL, HR, dst, src, A, B)
    allocate(v(n), hv(n), ref(n))
    call random_number(v)
    spHtimesV_p => spMatVec_gpu_main
    call spHtimesV_p(n, v, hv)
    ! reference loops (spMatVec_direct_main, src/DMRG_SUPERBLOCK_SETUP.f90:760-857)
    ref = 0d0
    do q = 1, nq
       do il = 1, Dls(q)
          do ir = 1, Drs(q)
             do jr = 1, Drs(q)
                ref(ir + (il-1)*Drs(q) + offs(q)) = ref(ir + (il-1)*Drs(q) + offs(q)) + &
                     HR(q)%m(ir, jr) * v(jr + (il-1)*Drs(q) + offs(q))
             enddo
             do jl = 1, Dls(q)
                ref(ir + (il-1)*Drs(q) + offs(q)) = ref(ir + (il-1)*Drs(q) + offs(q)) + &
                     HL(q)%m(il, jl) * v(ir + (jl-1)*Drs(q) + offs(q))
             enddo
          enddo
       enddo
    enddo
    do t = 1, nt
       allocate(C(size(B(t)%m, 1), size(A(t)%m, 2)))
       C = 0d0
       do aj = 1, size(A(t)%m, 2)
          do bi = 1, size(B(t)%m, 1)
             do bj = 1, size(B(t)%m, 2)
                C(bi, aj) = C(bi, aj) + B(t)%m(bi, bj) * v(bj + (aj-1)*size(B(t)%m, 2) + offs(src(t)))
             enddo
          enddo
       enddo
       do bi = 1, size(B(t)%m, 1)
          do ai = 1, size(A(t)%m, 1)
             do aj = 1, size(A(t)%m, 2)
                ref(bi + (ai-1)*size(B(t)%m, 1) + offs(dst(t))) = ref(bi + (ai-1)*size(B(t)%m, 1) + offs(dst(t))) + &
                     A(t)%m(ai, aj) * C(bi, aj)
             enddo
          enddo
       enddo
       deallocate(C)
    enddo
    res = maxval(abs(hv - ref)) / maxval(abs(ref))
    write(*, "(A,ES10.3)") "drop-in spMatVec_gpu_main vs direct loops, max rel err =", res
    if (res > 1d-13) then
       write(*, "(A)") "FAIL drop-in H*v"
       stop 1
    endif
    ! Lanczos on a symmetric superblock: keep only the sector-diagonal terms
    call nss_gpu_set_superblock(Dls, Drs, HL, HR, [integer ::], [integer ::], A(1:0), B(1:0))
    allocate(evecs(n, 1))
    call nss_gpu_lanczos(1, evals, evecs, tol=1d-12, seed=1, numop=nop, kmax=80)
    call spHtimesV_p(n, evecs(:, 1), hv)
    res = sqrt(sum((hv - evals(1) * evecs(:, 1))**2))
    write(*, "(A,F14.8,A,ES10.3,A,I4)") "drop-in Lanczos E0 =", evals(1), "  residual =", res, "  H*v calls =", nop
    if (res > 1d-8 * max(1d0, abs(evals(1)))) then
       write(*, "(A)") "FAIL drop-in Lanczos residual"
       stop 1
    endif
  end subroutine check_drop_in

end program spin_chain_golden
