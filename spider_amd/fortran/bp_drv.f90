! bp_drv.f90 -- `BP 3F` / `BP 32F` / `RB 3F` / `RB 32F` operation driver on libspbp.
!
! Keeps the reference operation's signature and answer sequence
! (bp3f.f:127-283, bp32f.f:102-410): the answers are read one per line from
! standard input, as a SPIDER procedure supplies them:
!
!   BP 3F :  template, file numbers, angles doc, symmetries doc ('*' = none), volume
!   BP 32F:  template, file numbers, angles doc, symmetries doc,
!            reconstructed volume, first sample volume, second sample volume
!   RB 3F / RB 32F: as BP 3F / BP 32F, then the template for the transformed
!            images ('*': not written; `rot***` or `rot@***`: every image after
!            its RTSQ, under the input's file numbers, bp32f.f:517-527,
!            826-827); registers 6, 7, 8 of the angles doc give the in-plane
!            angle and shifts applied by RTSQ before the insertion
!            (bp32f.f:805-837)
!
! Usage:  spbp_bp 3F|32F|RB3F|RB32F [data-extension] [FBS] < answers
! FBS: the RB pre-rotation by RTSF (Fourier-based bicubic interpolation), as
! after SPIDER's `MD` `FBS ON` (setmode.f:128,169; bp32f.f:811-816).
! Text after ';' on an answer line is a comment.  Images are SPIDER 2D files
! (`name***` templates, or `stack@***` over stacks and indexed stacks), the
! angles doc holds PSI, THETA, PHI
! in registers 1-3 per image number, volumes are written as SPIDER files.
program spbp_bp
   use iso_c_binding
   use spbp
   implicit none

   character(len=512) :: op, ext, template, numstr, angdoc, symdoc, outs(3), fname, rottmpl
   integer, allocatable :: nums(:)
   real(c_float), allocatable, target :: imgs(:, :, :), vol(:, :, :, :)
   real(c_double), allocatable :: angles(:, :)
   integer(c_int8_t), allocatable, target :: half(:)
   integer, allocatable :: dockeys(:), docn(:)
   real, allocatable :: docregs(:, :)
   real, allocatable :: img(:, :)
   type(spbp_opts) :: opts
   type(c_ptr) :: ctx
   integer :: nang, nx, ny, k, ierr, nout, kidx, i, j, iord
   integer(c_int) :: rc
   logical :: is32, isrb, fbs
   real(c_float), allocatable :: inplane(:, :)
   real(c_float), allocatable :: rotimgs(:, :, :)
   real(c_double), allocatable :: symang(:, :)
   integer :: nsym
   real :: x
   logical :: itmp

   call get_command_argument(1, op)
   ext = ' '
   if (command_argument_count() >= 2) call get_command_argument(2, ext)
   fbs = .false.
   if (command_argument_count() >= 3) then
      call get_command_argument(3, fname)
      fbs = trim(fname) == 'FBS' .or. trim(fname) == 'fbs'
   end if
   isrb = .false.
   select case (trim(op))
   case ('3F', '3f', 'BP 3F')
      is32 = .false.
      nout = 1
   case ('32F', '32f', 'BP 32F')
      is32 = .true.
      nout = 3
   case ('RB3F', 'rb3f', 'RB 3F')
      is32 = .false.
      isrb = .true.
      nout = 1
   case ('RB32F', 'rb32f', 'RB 32F')
      is32 = .true.
      isrb = .true.
      nout = 3
   case default
      write (*, '(A)') ' usage: spbp_bp 3F|32F|RB3F|RB32F [data-extension] [FBS] < answers'
      stop 2
   end select

   call answer('TEMPLATE FOR IMAGE FILES', template)
   call answer('FILE NUMBERS', numstr)
   call answer('ANGLES DOC', angdoc)
   call answer('SYMMETRIES DOC', symdoc)
   do k = 1, nout
      call answer('OUTPUT VOLUME', outs(k))
   end do
   rottmpl = '*'
   if (isrb) call answer('TRANSFORMED OUTPUT IMAGES TEMPLATE', rottmpl)

   call expand_numbers(numstr, nums)
   nang = size(nums)
   if (nang < 1) call errt(203, 'EMPTY FILE NUMBER LIST')
   write (*, '(A,I7)') '  NUMBER OF IMAGES: ', nang

   ! images
   do k = 1, nang
      call image_name(template, nums(k), ext, fname, kidx)
      call read_spider_2d(fname, kidx, img, ierr)
      if (ierr /= 0) call errt(ierr, 'CANNOT READ IMAGE: '//trim(fname))
      if (k == 1) then
         nx = size(img, 1)
         ny = size(img, 2)
         if (nx /= ny) call errt(101, 'ONLY WORKS ON SQUARE IMAGES')
         allocate (imgs(nx, ny, nang))
      else if (size(img, 1) /= nx .or. size(img, 2) /= ny) then
         call errt(101, 'IMAGES DIFFER IN SIZE: '//trim(fname))
      end if
      imgs(:, :, k) = img
   end do

   ! angles (BUILDM1: registers PSI, THETA, PHI of each image's key); RB:
   ! registers 6-8 (in-plane angle, x shift, y shift)
   call read_doc(angdoc, dockeys, docregs, docn)
   allocate (angles(3, nang))
   if (isrb) allocate (inplane(3, nang))
   do k = 1, nang
      j = findloc(dockeys, nums(k), dim=1)
      if (j == 0) call errt(102, 'MISSING ANGLE FOR IMAGE')
      angles(:, k) = real(docregs(1:3, j), c_double)
      if (isrb) then
         if (docn(j) < 8) call errt(102, 'MISSING ANGLE FOR IMAGE (REGISTERS 6-8)')
         inplane(:, k) = real(docregs(6:8, j), c_float)
      end if
   end do

   ! symmetries (BUILDS: keys 1..max, each row PSI, THETA, PHI)
   nsym = 0
   if (trim(symdoc) /= '*') then
      deallocate (dockeys, docregs, docn)
      call read_doc(symdoc, dockeys, docregs, docn)
      nsym = maxval(dockeys)
      allocate (symang(3, nsym))
      do k = 1, nsym
         j = findloc(dockeys, k, dim=1)
         if (j == 0) call errt(102, 'MISSING SYMMETRY')
         symang(:, k) = real(docregs(1:3, j), c_double)
      end do
   end if

   ! half-set split for BP 32F as RANDLIST (bp32f.f:544-553): first half
   ! true, then NANG random swaps; true -> first sample volume
   allocate (half(nang))
   if (is32) then
      half(1:nang / 2) = 0_c_int8_t
      half(nang / 2 + 1:nang) = 1_c_int8_t
      call random_seed()
      do k = 1, nang
         call random_number(x)
         iord = min(nang, max(1, int(x * nang + 0.5)))
         itmp = half(iord) == 0
         half(iord) = half(k)
         half(k) = merge(0_c_int8_t, 1_c_int8_t, itmp)
      end do
   end if

   opts%nx = nx
   opts%flags = ior(merge(SPBP_TWO_HALVES, 0, is32), merge(SPBP_FBS, 0, fbs))
   opts%device = -1
   rc = spbp_create(opts, ctx)
   if (rc /= SPBP_OK) call errt(int(rc), trim(spbp_errmsg(c_null_ptr)))
   if (nsym > 1) then
      rc = spbp_set_symmetries(ctx, int(nsym, c_int), symang)
      if (rc /= SPBP_OK) call errt(int(rc), trim(spbp_errmsg(ctx)))
   end if
   if (isrb .and. trim(rottmpl) /= '*') then
      ! the transformed images are written, then inserted as they are
      allocate (rotimgs(nx, nx, nang))
      if (fbs) then
         rc = spbp_rtsf_host(imgs, int(nang, c_int), int(nx, c_int), inplane, rotimgs)
      else
         rc = spbp_rtsq_host(imgs, int(nang, c_int), int(nx, c_int), inplane, rotimgs)
      end if
      if (rc /= SPBP_OK) call errt(int(rc), trim(spbp_errmsg(c_null_ptr)))
      call write_images(rottmpl, nums, ext, rotimgs)
      imgs = rotimgs
      deallocate (rotimgs)
      if (is32) then
         rc = spbp_insert_host(ctx, imgs, SPBP_F32, int(nang, c_int), angles, c_loc(half), c_null_ptr)
      else
         rc = spbp_insert_host(ctx, imgs, SPBP_F32, int(nang, c_int), angles, c_null_ptr, c_null_ptr)
      end if
   else if (isrb .and. is32) then
      rc = spbp_insert_rb_host(ctx, imgs, SPBP_F32, int(nang, c_int), angles, inplane, c_loc(half), c_null_ptr)
   else if (isrb) then
      rc = spbp_insert_rb_host(ctx, imgs, SPBP_F32, int(nang, c_int), angles, inplane, c_null_ptr, c_null_ptr)
   else if (is32) then
      rc = spbp_insert_host(ctx, imgs, SPBP_F32, int(nang, c_int), angles, c_loc(half), c_null_ptr)
   else
      rc = spbp_insert_host(ctx, imgs, SPBP_F32, int(nang, c_int), angles, c_null_ptr, c_null_ptr)
   end if
   if (rc /= SPBP_OK) call errt(int(rc), trim(spbp_errmsg(ctx)))

   allocate (vol(nx, nx, nx, nout))
   if (is32) then
      rc = spbp_finalize_host(ctx, c_loc(vol(1, 1, 1, 1)), c_loc(vol(1, 1, 1, 2)), c_loc(vol(1, 1, 1, 3)), c_null_ptr)
   else
      rc = spbp_finalize_host(ctx, c_loc(vol(1, 1, 1, 1)), c_null_ptr, c_null_ptr, c_null_ptr)
   end if
   if (rc /= SPBP_OK) call errt(int(rc), trim(spbp_errmsg(ctx)))
   rc = spbp_destroy(ctx)

   do k = 1, nout
      fname = with_ext(outs(k), ext)
      call write_spider_vol(fname, vol(:, :, :, k), ierr)
      if (ierr /= 0) call errt(ierr, 'CANNOT WRITE VOLUME: '//trim(fname))
      write (*, '(A,A,2(1X,ES13.5))') '  WROTE: ', trim(fname), minval(vol(:, :, :, k)), maxval(vol(:, :, :, k))
   end do

contains

   subroutine errt(code, msg)
      integer, intent(in) :: code
      character(len=*), intent(in) :: msg
      write (*, '(A,I0,A,A)') ' *** ERROR (', code, '): ', trim(msg)
      error stop 1
   end subroutine errt

   ! One answer line; text from ';' on is a comment.
   subroutine answer(prompt, val)
      character(len=*), intent(in) :: prompt
      character(len=*), intent(out) :: val
      character(len=1024) :: line
      integer :: ios, p
      read (*, '(A)', iostat=ios) line
      if (ios /= 0) call errt(203, 'NO ANSWER FOR: '//prompt)
      p = index(line, ';')
      if (p > 0) line = line(1:p - 1)
      val = adjustl(line)
   end subroutine answer

   ! "1-12", "1,3,5-9" or blank separated.
   subroutine expand_numbers(s, out)
      character(len=*), intent(in) :: s
      integer, allocatable, intent(out) :: out(:)
      character(len=len(s)) :: t
      integer :: p, q, a, b, n, ios, i
      integer :: buf(100000)
      t = s
      do i = 1, len(t)
         if (t(i:i) == ',') t(i:i) = ' '
      end do
      n = 0
      p = 1
      do while (p <= len_trim(t))
         do while (p <= len_trim(t) .and. t(p:p) == ' ')
            p = p + 1
         end do
         if (p > len_trim(t)) exit
         q = p
         do while (q <= len_trim(t) .and. t(q:q) /= ' ')
            q = q + 1
         end do
         i = index(t(p:q - 1), '-')
         if (i > 1) then
            read (t(p:p + i - 2), *, iostat=ios) a
            read (t(p + i:q - 1), *, iostat=ios) b
         else
            read (t(p:q - 1), *, iostat=ios) a
            b = a
         end if
         if (ios /= 0) call errt(203, 'BAD FILE NUMBER LIST: '//trim(s))
         do i = a, b
            n = n + 1
            if (n > size(buf)) call errt(203, 'FILE NUMBER LIST TOO LONG')
            buf(n) = i
         end do
         p = q
      end do
      allocate (out(n))
      out = buf(1:n)
   end subroutine expand_numbers

   function with_ext(name, ext) result(r)
      character(len=*), intent(in) :: name, ext
      character(len=512) :: r
      r = name
      if (len_trim(ext) > 0 .and. index(name, '.', back=.true.) <= index(name, '/', back=.true.)) &
         r = trim(name)//'.'//trim(ext)
   end function with_ext

   ! `name***` -> name001.ext ; `stack@***` -> stack.ext, image number k
   subroutine image_name(tmpl, k, ext, name, kidx)
      character(len=*), intent(in) :: tmpl, ext
      integer, intent(in) :: k
      character(len=*), intent(out) :: name
      integer, intent(out) :: kidx
      integer :: a, b, nst
      character(len=32) :: fmt, digits
      a = index(tmpl, '@')
      if (a > 0) then
         name = with_ext(tmpl(1:a - 1), ext)
         kidx = k
         return
      end if
      kidx = 0
      a = index(tmpl, '*')
      if (a == 0) call errt(203, 'TEMPLATE HAS NO *** FIELD: '//trim(tmpl))
      b = a
      do while (b < len_trim(tmpl) .and. tmpl(b + 1:b + 1) == '*')
         b = b + 1
      end do
      nst = b - a + 1
      write (fmt, '(A,I0,A,I0,A)') '(I', nst, '.', nst, ')'
      write (digits, fmt) k
      name = with_ext(tmpl(1:a - 1)//trim(digits)//tmpl(b + 1:len_trim(tmpl)), ext)
   end subroutine image_name

   elemental function bswap(x) result(y)
      real, intent(in) :: x
      real :: y
      integer :: i, j
      i = transfer(x, i)
      j = ior(ior(ishft(iand(i, 255), 24), ishft(iand(ishft(i, -8), 255), 16)), &
              ior(ishft(iand(ishft(i, -16), 255), 8), iand(ishft(i, -24), 255)))
      y = transfer(j, y)
   end function bswap

   elemental function bswapi(i) result(j)
      integer, intent(in) :: i
      integer :: j
      j = ior(ior(ishft(iand(i, 255), 24), ishft(iand(ishft(i, -8), 255), 16)), &
              ior(ishft(iand(ishft(i, -16), 255), 8), iand(ishft(i, -24), 255)))
   end function bswapi

   ! SPIDER 2D image (docs/image_doc.html): simple file, image kidx of a
   ! stack, or of an indexed stack (ISTACK = -MAXINDX; INDXREC =
   ! ceil(MAXINDX / NX) records of NX integers after the overall header give
   ! each image number's storage slot, 0 if absent; lunsethdr.f LUNREDINDX).
   subroutine read_spider_2d(fname, kidx, img, ierr)
      character(len=*), intent(in) :: fname
      integer, intent(in) :: kidx
      real, allocatable, intent(out) :: img(:, :)
      integer, intent(out) :: ierr
      real :: h(26)
      integer :: u, ios, nx, ny, nz, labbyt, istack, maxim, slot, indxrec
      integer(8) :: pos
      logical :: swap
      ierr = 0
      open (newunit=u, file=fname, access='stream', form='unformatted', status='old', action='read', iostat=ios)
      if (ios /= 0) then
         ierr = 18
         return
      end if
      read (u, pos=1, iostat=ios) h
      swap = .not. (h(12) >= 1.0 .and. h(23) == 4.0 * h(12))
      if (swap) h = bswap(h)
      nx = nint(h(12)); ny = nint(h(2)); nz = max(1, nint(h(1)))
      labbyt = nint(h(22)); istack = nint(h(24)); maxim = nint(h(26))
      if (ios /= 0 .or. nx < 1 .or. ny < 1 .or. nz /= 1 .or. labbyt < 1024) then
         ierr = 101
         close (u)
         return
      end if
      pos = int(labbyt, 8) + 1
      if (istack > 0) then
         if (kidx < 1 .or. kidx > maxim) then
            ierr = 102
            close (u)
            return
         end if
         pos = pos + int(kidx - 1, 8) * (int(labbyt, 8) + 4_8 * nx * ny) + labbyt
      else if (istack < 0) then
         if (kidx < 1 .or. kidx > -istack) then
            ierr = 102
            close (u)
            return
         end if
         read (u, pos=pos + 4_8 * (kidx - 1), iostat=ios) slot
         if (swap) slot = bswapi(slot)
         if (ios /= 0 .or. slot < 1) then
            ierr = 102
            close (u)
            return
         end if
         indxrec = (-istack + nx - 1) / nx
         pos = pos + int(indxrec, 8) * 4 * nx + int(slot - 1, 8) * (int(labbyt, 8) + 4_8 * nx * ny) + labbyt
      end if
      allocate (img(nx, ny))
      read (u, pos=pos, iostat=ios) img
      if (ios /= 0) ierr = 18
      if (swap) img = bswap(img)
      close (u)
   end subroutine read_spider_2d

   ! SPIDER header words 1,2,3,5-10,12,13,22-24,26,27 for an NX x NY x NZ
   ! image (with_stats: IMAMI = 1 and FMAX, FMIN, AV, SIG of v).
   subroutine spider_header(h, nx, ny, nz, v, with_stats, istack, maxim, imgnum, irec)
      real, allocatable, intent(out) :: h(:)
      integer, intent(in) :: nx, ny, nz, istack, maxim, imgnum, irec
      real, intent(in) :: v(:)
      logical, intent(in) :: with_stats
      integer :: labrec, labbyt
      double precision :: av
      labrec = 1024 / (4 * nx)
      if (mod(1024, 4 * nx) /= 0) labrec = labrec + 1
      labbyt = labrec * 4 * nx
      allocate (h(labbyt / 4))
      h = 0.0
      h(1) = nz; h(2) = ny; h(3) = irec; h(5) = merge(3.0, 1.0, nz > 1)
      if (with_stats) then
         av = sum(dble(v)) / size(v)
         h(6) = 1.0; h(7) = maxval(v); h(8) = minval(v); h(9) = real(av)
         h(10) = real(sqrt(max(0d0, sum((dble(v) - av)**2) / size(v))))
      end if
      h(12) = nx; h(13) = labrec; h(22) = labbyt; h(23) = 4 * nx
      h(24) = istack; h(26) = maxim; h(27) = imgnum
   end subroutine spider_header

   ! Images under a template, one per file number: `name***` -> one simple
   ! image each; `stk@***` -> one stack, image k = the k-th listed number
   ! (MAXIM = largest number; slots no number names hold IMGNUM 0).
   subroutine write_images(tmpl, nums, ext, v)
      character(len=*), intent(in) :: tmpl, ext
      integer, intent(in) :: nums(:)
      real, intent(in) :: v(:, :, :)
      character(len=512) :: name
      real, allocatable :: h(:), zero(:, :)
      integer :: u, ios, k, j, kidx, nx, ny, maxim, labrec
      nx = size(v, 1); ny = size(v, 2)
      if (index(tmpl, '@') == 0) then
         do k = 1, size(nums)
            call image_name(tmpl, nums(k), ext, name, kidx)
            call spider_header(h, nx, ny, 1, reshape(v(:, :, k), [nx * ny]), .true., 0, 0, 0, 0)
            h(3) = h(13) + ny
            open (newunit=u, file=name, access='stream', form='unformatted', status='replace', action='write', &
                  iostat=ios)
            if (ios /= 0) call errt(18, 'CANNOT WRITE IMAGE: '//trim(name))
            write (u, iostat=ios) h, v(:, :, k)
            if (ios /= 0) call errt(18, 'CANNOT WRITE IMAGE: '//trim(name))
            close (u)
         end do
         return
      end if
      if (minval(nums) < 1) call errt(203, 'STACK IMAGE NUMBERS START AT 1')
      call image_name(tmpl, nums(1), ext, name, kidx)
      maxim = maxval(nums)
      allocate (zero(nx, ny))
      zero = 0.0
      call spider_header(h, nx, ny, 1, [0.0], .false., 2, maxim, 0, 0)
      labrec = nint(h(13))
      h(3) = labrec + ny
      open (newunit=u, file=name, access='stream', form='unformatted', status='replace', action='write', iostat=ios)
      if (ios /= 0) call errt(18, 'CANNOT WRITE IMAGE STACK: '//trim(name))
      write (u, iostat=ios) h
      do k = 1, maxim
         j = findloc(nums, k, dim=1)
         if (j > 0) then
            call spider_header(h, nx, ny, 1, reshape(v(:, :, j), [nx * ny]), .true., 0, 0, k, labrec + ny)
            write (u, iostat=ios) h, v(:, :, j)
         else
            call spider_header(h, nx, ny, 1, [0.0], .false., 0, 0, 0, labrec + ny)
            write (u, iostat=ios) h, zero
         end if
         if (ios /= 0) call errt(18, 'CANNOT WRITE IMAGE STACK: '//trim(name))
      end do
      close (u)
   end subroutine write_images

   ! Simple SPIDER volume, machine byte order; header words 1,2,3,5-10,12,13,22,23.
   subroutine write_spider_vol(fname, v, ierr)
      character(len=*), intent(in) :: fname
      real, intent(in) :: v(:, :, :)
      integer, intent(out) :: ierr
      real, allocatable :: h(:)
      integer :: u, ios, nx, ny, nz, labrec, labbyt
      double precision :: av
      nx = size(v, 1); ny = size(v, 2); nz = size(v, 3)
      labrec = 1024 / (4 * nx)
      if (mod(1024, 4 * nx) /= 0) labrec = labrec + 1
      labbyt = labrec * 4 * nx
      allocate (h(labbyt / 4))
      h = 0.0
      av = sum(dble(v)) / size(v)
      h(1) = nz; h(2) = ny; h(3) = labrec + ny * nz; h(5) = merge(3.0, 1.0, nz > 1)
      h(6) = 1.0; h(7) = maxval(v); h(8) = minval(v); h(9) = real(av)
      h(10) = real(sqrt(max(0d0, sum((dble(v) - av)**2) / size(v))))
      h(12) = nx; h(13) = labrec; h(22) = labbyt; h(23) = 4 * nx
      ierr = 0
      open (newunit=u, file=fname, access='stream', form='unformatted', status='replace', action='write', iostat=ios)
      if (ios /= 0) then
         ierr = 18
         return
      end if
      write (u, iostat=ios) h, v
      if (ios /= 0) ierr = 18
      close (u)
   end subroutine write_spider_vol

   ! SPIDER document file: key, register count, registers; ';' lines are comments.
   ! keys, their first 8 registers (zero beyond the line's count) and counts
   subroutine read_doc(fname, keys, regs, nregs)
      character(len=*), intent(in) :: fname
      integer, allocatable, intent(out) :: keys(:)
      real, allocatable, intent(out) :: regs(:, :)
      integer, allocatable, intent(out) :: nregs(:)
      character(len=2048) :: line, tline
      integer :: u, ios, key, n, m, i
      real :: vals(64)
      integer, allocatable :: tk(:), tn(:)
      real, allocatable :: tr(:, :)
      open (newunit=u, file=with_ext(fname, ext), status='old', action='read', iostat=ios)
      if (ios /= 0) call errt(18, 'CANNOT OPEN DOC FILE: '//trim(fname))
      allocate (keys(0), regs(8, 0), nregs(0))
      m = 0
      do
         read (u, '(A)', iostat=ios) line
         if (ios /= 0) exit
         if (len_trim(line) == 0) cycle
         tline = adjustl(line)
         if (line(2:2) == ';' .or. tline(1:1) == ';') cycle
         vals = 0.0
         read (line, *, iostat=ios) key, n
         if (ios == 0 .and. n >= 1 .and. n <= 64) read (line, *, iostat=ios) key, n, (vals(i), i=1, n)
         if (ios /= 0) then
            read (line, '(I6,I1,9G12.3)', iostat=ios) key, n, (vals(i), i=1, min(n, 9))
            if (ios /= 0) cycle
         end if
         if (n < 3) cycle
         m = m + 1
         allocate (tk(m), tr(8, m), tn(m))
         tk(1:m - 1) = keys
         tr(:, 1:m - 1) = regs
         tn(1:m - 1) = nregs
         tk(m) = key
         tr(:, m) = vals(1:8)
         tn(m) = n
         call move_alloc(tk, keys)
         call move_alloc(tr, regs)
         call move_alloc(tn, nregs)
      end do
      close (u)
   end subroutine read_doc

end program spbp_bp
