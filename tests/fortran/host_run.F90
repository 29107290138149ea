! Test driver of the Fortran host module MCTQMC (w2dynamics_amd/fortran/ctqmc_host.F90).
! It reads the same input directory as the reference driver in oracle/ref_driver.F90
! (run.txt, params.txt, umatrix.bin, ftau.bin, muimp.bin), calls the solver in the
! order of w2dyn/dmft/cthyb_solver.py:240-342
!   init_paras -> init_rnd -> init_solver -> set_taudiffmax -> ctqmc_warmup ->
!   ctqmc_measure(1,1)
! and writes <dir>/results.bin in the reference driver's layout, so that the two
! programs' outputs compare byte layout for byte layout.
!
! run.txt: mode (must be "run"), nb, nft, seed, nwarm, nmoves, taudiffmax
! [, n_chains]
program host_run
   use iso_c_binding
   use MCTQMC
   implicit none
   character(len=4096) :: dir
   character(len=16) :: mode
   integer :: nb, nft, seed, ios, nch
   integer(c_int64_t) :: nwarm, nmoves
   real(c_double) :: taudiffmax, tdm
   real(c_double), allocatable :: umat(:,:,:,:), ftau(:,:,:,:,:), muimp(:,:,:,:), screen(:,:,:,:,:)
   real(c_double), allocatable :: apt0(:)
   character(len=:), allocatable :: pstr
   character(len=600) :: line

   call get_command_argument(1, dir)
   open (10, file=trim(dir)//'/run.txt', status='old', action='read')
   read (10, *) mode
   read (10, *) nb
   read (10, *) nft
   read (10, *) seed
   read (10, *) nwarm
   read (10, *) nmoves
   read (10, *) taudiffmax
   read (10, *, iostat=ios) nch
   if (ios == 0) n_chains = nch
   close (10)
   if (trim(mode) /= 'run' .and. trim(mode) /= 'calib') stop 'host_run: modes run and calib'

   pstr = ''
   open (11, file=trim(dir)//'/params.txt', status='old', action='read')
   do
      read (11, '(A)', iostat=ios) line
      if (ios /= 0) exit
      pstr = pstr//trim(line)//char(10)
   end do
   close (11)

   allocate (umat(2*nb, 2*nb, 2*nb, 2*nb), ftau(nb, 2, nb, 2, nft), screen(nb, 2, nb, 2, nft))
   allocate (muimp(nb, 2, nb, 2))
   screen = 0
   call read_array(trim(dir)//'/umatrix.bin', umat, size(umat))
   call read_array(trim(dir)//'/ftau.bin', ftau, size(ftau))
   call read_array(trim(dir)//'/muimp.bin', muimp, size(muimp))

   call init_paras(pstr)
   call init_rnd(seed)
   call init_solver(umat, ftau, muimp, screen, 1, nft, nb)
   if (aborted /= 0) stop 3
   if (taudiffmax > 0) call set_taudiffmax(taudiffmax)
   if (trim(mode) == 'calib') then
      ! cthyb_solver.py:899-944 up to the phase-2 histogram (the driver's fit
      ! of it is left out: the run keeps the phase-1 TaudiffMax)
      call init_counters()
      call ctqmc_calibrate(.true., 1000, 100, 1000, 100000)
      if (aborted /= 0) stop 5
      apt0 = AccPairTau          ! phase 2 deallocates AccPairTau, as the reference does
      tdm = AccPairTau(int(0.995d0*1000) + 1)
      call set_taudiffmax(tdm)
      call init_counters()
      call ctqmc_calibrate(.false., 1000, 100, 1000, 100000)
      if (aborted /= 0) stop 6
      open (71, file=trim(dir)//'/calib.bin', access='stream', form='unformatted', status='replace', &
            action='write')
      write (71) int(size(apt0), c_int32_t)
      write (71) apt0
      write (71) tdm
      write (71) int(size(AccPair, 3), c_int32_t)
      write (71) AccPair
   end if
   call ctqmc_warmup(nwarm)
   call ctqmc_measure(1, 1)
   if (aborted /= 0) stop 4
   if (trim(mode) == 'calib') then
      write (71) int(size(AccPair, 3), c_int32_t)
      write (71) AccPair
      write (71) OuterStateHisto, SignHistoSuperstates, SignHistoStates, TraceContribSuperstates, &
                 TraceContribStates
      write (71) Gtau_mean_step, Gtau_mid_step, sign_step
      write (71) rho1, rho2
      close (71)
   end if

   open (70, file=trim(dir)//'/results.bin', access='stream', form='unformatted', status='replace', &
         action='write')
   write (70) mean_sign
   write (70) Gtau
   write (70) GLeg
   write (70) occ
   write (70) DensityMatrix
   write (70) Histo
   write (70) AccAdd(1), AccRem(1), AccGlob, AccShift
   write (70) OuterSuperstateHisto
   close (70)
   call dest_ctqmc()
   call dest_paras()

contains

   subroutine read_array(fname, arr, n)
      character(*), intent(in) :: fname
      integer, intent(in) :: n
      real(c_double), intent(inout) :: arr(n)
      open (20, file=fname, access='stream', form='unformatted', status='old', action='read')
      read (20) arr
      close (20)
   end subroutine read_array
end program host_run
