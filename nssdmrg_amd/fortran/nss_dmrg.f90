!> Fortran host binding of the MI355X superblock path (include/nssdmrg.h).
!>
!> What a maintainer of the reference adds to use the GPU path:
!>  * spMatVec_gpu_main conforms to the abstract interface sparse_HxV
!>    (src/DMRG_GLOBAL.f90:41-53) and can be bound to spHtimesV_p in place of
!>    spMatVec_direct_main (src/DMRG_SUPERBLOCK.f90:556-560);
!>  * nss_gpu_set_superblock uploads the per-sector operator blocks that
!>    Setup_SuperBlock_Direct builds (src/DMRG_SUPERBLOCK_SETUP.f90:113-693);
!>  * nss_gpu_lanczos replaces sp_eigh in sb_diag (src/DMRG_SUPERBLOCK.f90:412-425);
!>  * init_dmrg_gpu / run_dmrg_gpu / finalize_dmrg_gpu drive the whole
!>    device-resident DMRG (init_dmrg / run_DMRG / finalize_dmrg,
!>    src/DMRG_MAIN.f90:22-132).
!> Every C call returns a status; a nonzero status stops with the library's
!> message, as the reference stops on errors.
module nss_dmrg
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: NSS_OK = 0
  integer(c_int), parameter, public :: NSS_FP64 = 0, NSS_FP32_MIXED = 1
  integer(c_int), parameter, public :: NSS_ROW_MAJOR = 0, NSS_COL_MAJOR = 1
  integer(c_int), parameter, public :: NSS_MODEL_SPIN12 = 0, NSS_MODEL_SPIN1 = 1, NSS_MODEL_HUBBARD = 2

  type, bind(C), public :: nss_model
     integer(c_int32_t) :: model = 0, hfmode = 1
     real(c_double) :: hij(4) = 0d0
     real(c_double) :: uloc = 0d0
     real(c_double) :: hvec(2) = 0d0
     real(c_double) :: eloc(2) = 0d0
  end type nss_model

  type, bind(C), public :: nss_params
     integer(c_int32_t) :: ldmrg = 5, mdmrg = 0
     real(c_double) :: edmrg = 0d0
     integer(c_int32_t) :: nsweep = 0
     type(c_ptr) :: msweep = c_null_ptr, esweep = c_null_ptr
     real(c_double) :: target_qn(2) = 0d0
     integer(c_int32_t) :: neigen = 2, lanc_dim_threshold = 1024, lanc_kmax = 150, lanc_maxiter = 512
     real(c_double) :: lanc_tol = 1d-12
     integer(c_int64_t) :: seed = 0
     integer(c_int32_t) :: v0_mode = 0
     real(c_double) :: deg_thr = 0.1d0
     integer(c_int32_t) :: precision = 0, verbose = 0, predict = 0
     integer(c_int32_t) :: pbc = 0
     integer(c_int32_t) :: qn_type = 0
     real(c_double) :: qn_trunc_error = 0d0
     integer(c_int32_t) :: qn_trunc_dim = 1
  end type nss_params

  type, bind(C), public :: nss_step_info
     integer(c_int32_t) :: type, label, sweep, left_len, right_len, L, neigen
     real(c_double) :: energies(4)
     real(c_double) :: entropy_left, entropy_right
     integer(c_int32_t) :: m_left, m_right, renormalized
     real(c_double) :: trunc_left, trunc_right, cut_left, cut_right
     integer(c_int64_t) :: m_sb
     integer(c_int32_t) :: nsect, nterms, numop, numop_fp32
     real(c_double) :: hv_flops
     real(c_double) :: t_step, t_diag, t_hv, t_rdm, t_renorm, t_enlarge
     integer(c_int32_t) :: link
     real(c_double) :: sn_left, sc_left, sn_right, sc_right
  end type nss_step_info

  !> dense operator block of a superblock sector (column-major, Fortran)
  type, public :: dense_block
     real(c_double), allocatable :: m(:,:)
  end type dense_block

  interface
     function c_nss_last_error() bind(C, name="nss_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function c_nss_last_error
     function c_nss_ctx_create(device, out) bind(C, name="nss_ctx_create") result(r)
       import :: c_int, c_ptr
       integer(c_int), value :: device
       type(c_ptr) :: out
       integer(c_int) :: r
     end function c_nss_ctx_create
     function c_nss_ctx_destroy(ctx) bind(C, name="nss_ctx_destroy") result(r)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: r
     end function c_nss_ctx_destroy
     function c_nss_sb_set_layout(ctx, nq, dl, dr, off) bind(C, name="nss_sb_set_layout") result(r)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: nq
       integer(c_int32_t) :: dl(*), dr(*)
       type(c_ptr), value :: off
       integer(c_int) :: r
     end function c_nss_sb_set_layout
     function c_nss_sb_upload_blocks(ctx, hl, hr, nterms, dst, src, a, b, layout, precision) &
          bind(C, name="nss_sb_upload_blocks") result(r)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       type(c_ptr) :: hl(*), hr(*), a(*), b(*)
       integer(c_int), value :: nterms, layout, precision
       integer(c_int32_t) :: dst(*), src(*)
       integer(c_int) :: r
     end function c_nss_sb_upload_blocks
     function c_nss_sb_hv(ctx, v, hv, n) bind(C, name="nss_sb_hv") result(r)
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double) :: v(*), hv(*)
       integer(c_int64_t), value :: n
       integer(c_int) :: r
     end function c_nss_sb_hv
     function c_nss_lanczos(ctx, neigen, kmax, maxiter, tol, seed, v0_mode, precision, evals, evecs, numop) &
          bind(C, name="nss_lanczos") result(r)
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: neigen, kmax, maxiter, v0_mode, precision
       real(c_double), value :: tol
       integer(c_int64_t), value :: seed
       real(c_double) :: evals(*), evecs(*)
       integer(c_int) :: numop
       integer(c_int) :: r
     end function c_nss_lanczos
     function c_nss_dmrg_create(ctx, model, params, out) bind(C, name="nss_dmrg_create") result(r)
       import :: c_int, c_ptr, nss_model, nss_params
       type(c_ptr), value :: ctx
       type(nss_model) :: model
       type(nss_params) :: params
       type(c_ptr) :: out
       integer(c_int) :: r
     end function c_nss_dmrg_create
     function c_nss_dmrg_run(d, type) bind(C, name="nss_dmrg_run") result(r)
       import :: c_int, c_ptr
       type(c_ptr), value :: d
       integer(c_int), value :: type
       integer(c_int) :: r
     end function c_nss_dmrg_run
     function c_nss_dmrg_nsteps(d) bind(C, name="nss_dmrg_nsteps") result(r)
       import :: c_int, c_ptr
       type(c_ptr), value :: d
       integer(c_int) :: r
     end function c_nss_dmrg_nsteps
     function c_nss_dmrg_step_info(d, i, out) bind(C, name="nss_dmrg_step_info") result(r)
       import :: c_int, c_ptr, nss_step_info
       type(c_ptr), value :: d
       integer(c_int), value :: i
       type(nss_step_info) :: out
       integer(c_int) :: r
     end function c_nss_dmrg_step_info
     function c_nss_dmrg_destroy(d) bind(C, name="nss_dmrg_destroy") result(r)
       import :: c_int, c_ptr
       type(c_ptr), value :: d
       integer(c_int) :: r
     end function c_nss_dmrg_destroy
     function c_nss_dmrg_measure(d, nops, ops, npos, pos, out) bind(C, name="nss_dmrg_measure") result(r)
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: d
       integer(c_int), value :: nops, npos
       real(c_double), intent(in) :: ops(*)
       integer(c_int32_t), intent(in) :: pos(*)
       real(c_double), intent(out) :: out(*)
       integer(c_int) :: r
     end function c_nss_dmrg_measure
     function c_strlen(s) bind(C, name="strlen") result(n)
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: n
     end function c_strlen
  end interface

  type(c_ptr), save :: ctx_ = c_null_ptr
  type(c_ptr), save :: dmrg_ = c_null_ptr
  integer(c_int32_t), allocatable, target, save :: msweep_(:)
  real(c_double), allocatable, target, save :: esweep_(:)

  public :: nss_gpu_init, nss_gpu_finalize, nss_gpu_set_superblock, spMatVec_gpu_main, nss_gpu_lanczos
  public :: init_dmrg_gpu, run_dmrg_gpu, finalize_dmrg_gpu, dmrg_gpu_nsteps, dmrg_gpu_step, dmrg_gpu_measure
  public :: nss_check

contains

  !> Message of the last failing C call.
  function last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n, i
    p = c_nss_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = s(i)
    enddo
  end function last_error

  subroutine nss_check(ierr, where)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: where
    if (ierr /= NSS_OK) then
       write(*, "(A)") trim(where)//": "//last_error()
       stop 1
    endif
  end subroutine nss_check

  subroutine nss_gpu_init(device)
    integer, intent(in), optional :: device
    integer(c_int) :: dev
    dev = 0
    if (present(device)) dev = device
    if (.not. c_associated(ctx_)) call nss_check(c_nss_ctx_create(dev, ctx_), "nss_ctx_create")
  end subroutine nss_gpu_init

  subroutine nss_gpu_finalize()
    if (c_associated(dmrg_)) call finalize_dmrg_gpu()
    if (c_associated(ctx_)) call nss_check(c_nss_ctx_destroy(ctx_), "nss_ctx_destroy")
    ctx_ = c_null_ptr
  end subroutine nss_gpu_finalize

  !> Upload the direct-mode superblock (Dls, Drs and the dense forms of
  !> Hleft(q), Hright(q), A(it,q), B(it,q) with row/col offsets given as
  !> destination/source sector indices, 1-based as in the reference).
  subroutine nss_gpu_set_superblock(Dls, Drs, Hleft, Hright, dst, src, A, B, precision)
    integer, intent(in) :: Dls(:), Drs(:), dst(:), src(:)
    type(dense_block), intent(in), target :: Hleft(:), Hright(:), A(:), B(:)
    integer, intent(in), optional :: precision
    integer(c_int32_t), allocatable :: dl(:), dr(:), d0(:), s0(:)
    type(c_ptr), allocatable :: phl(:), phr(:), pa(:), pb(:)
    integer :: q, t, nq, nt
    integer(c_int) :: prec
    nq = size(Dls)
    nt = size(dst)
    prec = NSS_FP64
    if (present(precision)) prec = precision
    allocate(dl(nq), dr(nq), d0(max(nt, 1)), s0(max(nt, 1)))
    dl = int(Dls, c_int32_t)
    dr = int(Drs, c_int32_t)
    allocate(phl(nq), phr(nq), pa(max(nt, 1)), pb(max(nt, 1)))
    do q = 1, nq
       phl(q) = c_loc(Hleft(q)%m)
       phr(q) = c_loc(Hright(q)%m)
    enddo
    do t = 1, nt
       d0(t) = int(dst(t) - 1, c_int32_t)
       s0(t) = int(src(t) - 1, c_int32_t)
       pa(t) = c_loc(A(t)%m)
       pb(t) = c_loc(B(t)%m)
    enddo
    call nss_check(c_nss_sb_set_layout(ctx_, int(nq, c_int), dl, dr, c_null_ptr), "nss_sb_set_layout")
    call nss_check(c_nss_sb_upload_blocks(ctx_, phl, phr, int(nt, c_int), d0, s0, pa, pb, NSS_COL_MAJOR, prec), &
         "nss_sb_upload_blocks")
  end subroutine nss_gpu_set_superblock

  !> H_sb . v on the GPU; conforms to sparse_HxV (src/DMRG_GLOBAL.f90:41-53).
  subroutine spMatVec_gpu_main(Nloc, v, Hv)
    integer :: Nloc
    real(8), dimension(Nloc) :: v
    real(8), dimension(Nloc) :: Hv
    call nss_check(c_nss_sb_hv(ctx_, v, Hv, int(Nloc, c_int64_t)), "spMatVec_gpu_main")
  end subroutine spMatVec_gpu_main

  !> Lowest Neigen eigenpairs of the uploaded superblock (sp_eigh's role).
  subroutine nss_gpu_lanczos(Neigen, evals, evecs, tol, seed, numop, kmax, precision)
    integer, intent(in) :: Neigen
    real(8), intent(out) :: evals(Neigen)
    real(8), intent(out) :: evecs(:,:)
    real(8), intent(in), optional :: tol
    integer, intent(in), optional :: seed, kmax, precision
    integer, intent(out), optional :: numop
    real(c_double) :: t
    integer(c_int64_t) :: sd
    integer(c_int) :: nop, km, pr
    t = 1d-12
    if (present(tol)) t = tol
    sd = 0
    if (present(seed)) sd = seed
    km = 150
    if (present(kmax)) km = kmax
    pr = NSS_FP64
    if (present(precision)) pr = precision
    call nss_check(c_nss_lanczos(ctx_, int(Neigen, c_int), km, 512_c_int, t, sd, 0_c_int, pr, evals, evecs, nop), &
         "nss_lanczos")
    if (present(numop)) numop = nop
  end subroutine nss_gpu_lanczos

  !> init_dmrg(Hij, ModelDot) on the device-resident engine.
  subroutine init_dmrg_gpu(Hij, model, par, msweep, esweep)
    real(8), intent(in) :: Hij(2, 2)
    type(nss_model), intent(inout) :: model
    type(nss_params), intent(inout) :: par
    integer, intent(in), optional :: msweep(:)
    real(8), intent(in), optional :: esweep(:)
    call nss_gpu_init()
    model%hij = [Hij(1, 1), Hij(1, 2), Hij(2, 1), Hij(2, 2)]
    if (present(msweep)) then
       allocate(msweep_(size(msweep)))
       msweep_ = int(msweep, c_int32_t)
       par%nsweep = size(msweep)
       par%msweep = c_loc(msweep_)
       if (present(esweep)) then
          allocate(esweep_(size(esweep)))
          esweep_ = esweep
          par%esweep = c_loc(esweep_)
       endif
    endif
    call nss_check(c_nss_dmrg_create(ctx_, model, par, dmrg_), "init_dmrg_gpu")
  end subroutine init_dmrg_gpu

  !> run_DMRG: type 'i' (infinite) or 'f' (finite).
  subroutine run_dmrg_gpu(type)
    character(len=1), intent(in) :: type
    call nss_check(c_nss_dmrg_run(dmrg_, int(ichar(type), c_int)), "run_dmrg_gpu")
  end subroutine run_dmrg_gpu

  integer function dmrg_gpu_nsteps()
    dmrg_gpu_nsteps = c_nss_dmrg_nsteps(dmrg_)
  end function dmrg_gpu_nsteps

  !> record of DMRG step i (1-based)
  function dmrg_gpu_step(i) result(info)
    integer, intent(in) :: i
    type(nss_step_info) :: info
    call nss_check(c_nss_dmrg_step_info(dmrg_, int(i - 1, c_int), info), "dmrg_gpu_step")
  end function dmrg_gpu_step

  !> Measure_DMRG(Op, pos, avOp) on the ground state of the last step:
  !> ops(:,:,k) site operators (site basis, QN conserving), pos 1-based chain
  !> positions; avOp(k, ipos) = <psi|O_k(pos(ipos))|psi>.
  subroutine dmrg_gpu_measure(ops, pos, avOp)
    real(8), intent(in) :: ops(:, :, :)
    integer, intent(in) :: pos(:)
    real(8), allocatable, intent(out) :: avOp(:, :)
    real(c_double), allocatable :: opsT(:), out(:)
    integer(c_int32_t), allocatable :: p(:)
    integer :: k, d, nops, np
    d = size(ops, 1)
    nops = size(ops, 3)
    np = size(pos)
    allocate(opsT(d * d * nops), out(nops * np), p(np))
    do k = 1, nops   ! row-major for the C side
       opsT((k - 1) * d * d + 1:k * d * d) = reshape(transpose(ops(:, :, k)), [d * d])
    enddo
    p = int(pos, c_int32_t)
    call nss_check(c_nss_dmrg_measure(dmrg_, int(nops, c_int), opsT, int(np, c_int), p, out), "dmrg_gpu_measure")
    allocate(avOp(nops, np))
    do k = 1, nops
       avOp(k, :) = out((k - 1) * np + 1:k * np)
    enddo
  end subroutine dmrg_gpu_measure

  subroutine finalize_dmrg_gpu()
    if (c_associated(dmrg_)) call nss_check(c_nss_dmrg_destroy(dmrg_), "finalize_dmrg_gpu")
    dmrg_ = c_null_ptr
    if (allocated(msweep_)) deallocate(msweep_)
    if (allocated(esweep_)) deallocate(esweep_)
  end subroutine finalize_dmrg_gpu

end module nss_dmrg
