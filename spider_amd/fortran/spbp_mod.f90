! spbp_mod.f90 -- ISO_C_BINDING interface to libspbp (include/spbp.h).
!
! Lets Fortran code call the MI355X backprojection engine the way SPIDER's
! BP32F / BP3F (bp32f.f:57, bp3f.f:79) call their engine BP32FQ / BP3FQ and
! the finalisation NRMW2 + WINDKB2A, but with image buffers instead of
! logical units.  Host-array entry points only; see spbp.h for the rest.
module spbp
   use iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: SPBP_OK = 0, SPBP_ERR_ALLOC = 46, SPBP_ERR_SIZE = 101, &
      SPBP_ERR_ANGLE = 102, SPBP_ERR_FFT = 201, SPBP_ERR_DEVICE = 202, SPBP_ERR_ARG = 203, &
      SPBP_ERR_NONFINITE = 205
   integer(c_int), parameter, public :: SPBP_F32 = 0, SPBP_BF16 = 1
   integer(c_int), parameter, public :: SPBP_TWO_HALVES = 1, SPBP_EXACT_PI = 2, SPBP_TWOSUM_W = 4, &
      SPBP_CHECK_FINITE = 8, SPBP_TIMING = 16, SPBP_GEOM_F64 = 32, SPBP_PLAIN_THRESHOLD = 64, SPBP_HIPFFT = 128, &
      SPBP_FBS = 256
   integer(c_int), parameter, public :: SPBP_STRAT_AUTO = 0, SPBP_STRAT_DIRECT = 1, SPBP_STRAT_BRICK = 2, &
      SPBP_STRAT_GATHER = 3

   type, bind(c), public :: spbp_opts
      integer(c_int) :: nx = 0
      integer(c_int) :: flags = 0
      integer(c_int) :: device = -1
      integer(c_int) :: batch = 0
      integer(c_int) :: strategy = 0
      integer(c_int) :: brick_proj = 0
   end type spbp_opts

   public :: spbp_create, spbp_destroy, spbp_insert_host, spbp_finalize_host, spbp_reset, spbp_errmsg, &
      spbp_set_symmetries, spbp_insert_rb_host, spbp_rtsq_host, spbp_rtsf_host

   interface
      integer(c_int) function spbp_create(opts, ctx) bind(c, name='spbp_create')
         import :: c_int, c_ptr, spbp_opts
         type(spbp_opts), intent(in) :: opts
         type(c_ptr), intent(out) :: ctx
      end function spbp_create

      integer(c_int) function spbp_destroy(ctx) bind(c, name='spbp_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function spbp_destroy

      integer(c_int) function spbp_reset(ctx, stream) bind(c, name='spbp_reset')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, stream
      end function spbp_reset

      ! imgs: NX x NX x nimg REAL (= PROJ(NX,NX) per image); angles: 3 x nimg
      ! DOUBLE PRECISION doc registers (PSI, THETA, PHI); half: c_loc of an
      ! INTEGER(c_int8_t) array of 0/1, or c_null_ptr.
      integer(c_int) function spbp_insert_host(ctx, imgs, dtype, nimg, angles, half, stream) &
            bind(c, name='spbp_insert_host')
         import :: c_int, c_ptr, c_float, c_double
         type(c_ptr), value :: ctx
         real(c_float), intent(in) :: imgs(*)
         integer(c_int), value :: dtype, nimg
         real(c_double), intent(in) :: angles(*)
         type(c_ptr), value :: half, stream
      end function spbp_insert_host

      ! RB 3F / RB 32F: as spbp_insert_host, each image first rotated and
      ! shifted by RTSQ; inplane: 3 x nimg REAL (doc registers 6, 7, 8)
      integer(c_int) function spbp_insert_rb_host(ctx, imgs, dtype, nimg, angles, inplane, half, stream) &
            bind(c, name='spbp_insert_rb_host')
         import :: c_int, c_ptr, c_float, c_double
         type(c_ptr), value :: ctx
         real(c_float), intent(in) :: imgs(*)
         integer(c_int), value :: dtype, nimg
         real(c_double), intent(in) :: angles(*)
         real(c_float), intent(in) :: inplane(*)
         type(c_ptr), value :: half, stream
      end function spbp_insert_rb_host

      ! RTSQ of host images (the RB transformed images); params 3 x nimg
      integer(c_int) function spbp_rtsq_host(imgs, nimg, nx, params, out) bind(c, name='spbp_rtsq_host')
         import :: c_int, c_float
         real(c_float), intent(in) :: imgs(*)
         integer(c_int), value :: nimg, nx
         real(c_float), intent(in) :: params(*)
         real(c_float), intent(out) :: out(*)
      end function spbp_rtsq_host

      ! RTSF of host images (FBS ON)
      integer(c_int) function spbp_rtsf_host(imgs, nimg, nx, params, out) bind(c, name='spbp_rtsf_host')
         import :: c_int, c_float
         real(c_float), intent(in) :: imgs(*)
         integer(c_int), value :: nimg, nx
         real(c_float), intent(in) :: params(*)
         real(c_float), intent(out) :: out(*)
      end function spbp_rtsf_host

      ! symmetries doc rows: 3 x nsym DOUBLE PRECISION (PSI, THETA, PHI)
      integer(c_int) function spbp_set_symmetries(ctx, nsym, angles) bind(c, name='spbp_set_symmetries')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nsym
         real(c_double), intent(in) :: angles(*)
      end function spbp_set_symmetries

      ! NX x NX x NX REAL volumes, c_loc(...) or c_null_ptr each
      integer(c_int) function spbp_finalize_host(ctx, out_full, out_h1, out_h2, stream) &
            bind(c, name='spbp_finalize_host')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, out_full, out_h1, out_h2, stream
      end function spbp_finalize_host

      type(c_ptr) function c_spbp_last_error(ctx) bind(c, name='spbp_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function c_spbp_last_error
   end interface

contains

   ! Message of the last failure on ctx (or of this thread when ctx is null).
   function spbp_errmsg(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=512) :: msg
      character(kind=c_char), pointer :: s(:)
      type(c_ptr) :: p
      integer :: i
      msg = ' '
      p = c_spbp_last_error(ctx)
      if (.not. c_associated(p)) return
      call c_f_pointer(p, s, [512])
      do i = 1, 512
         if (s(i) == c_null_char) exit
         msg(i:i) = s(i)
      end do
   end function spbp_errmsg

end module spbp
