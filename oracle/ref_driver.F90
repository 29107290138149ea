! Driver for the reference CT-HYB core (w2dynamics src/ctqmc_fortran), used only to
! produce golden vectors for the tests.  It is our own program: it links the
! reference's modules unchanged and calls their public procedures in the order the
! Python solver does (w2dyn/dmft/cthyb_solver.py:240-342):
!   init_paras -> init_rnd -> init_solver -> set_taudiffmax -> warmup/measure.
!
! Modes (first line of <dir>/run.txt):
!   log   : run NMOVES Metropolis steps with the same dispatch as ctqmc_warmup
!           (CTQMC.F90:8213-8254) and write one record per move to <dir>/moves.bin
!   run   : ctqmc_warmup(NWARM) + ctqmc_measure(1,1) and dump the observables
!   eval  : set_mc_config for each configuration in <dir>/configs.bin and dump the
!           trace / determinant / inverse hybridisation matrices
!   runcfg: after one line on standard input (or its end), set_mc_config with the
!           first configuration of <dir>/configs.bin (a thermalised state, as
!           ReuseMCConfig does), then as `run`; also
!           writes <dir>/timing.txt (time_sim of the sampling and the final order)
!
! run.txt layout (free format, one item per line):
!   mode, nb, nft, seed, nwarm, nmoves, taudiffmax
program ref_driver
   use iso_c_binding
   use MParameters
   use MStates
   use MTrace
   use MMatrixUpdate
   use MRandom
   use MCTQMC
   implicit none

   character(len=4096)          :: dir
   character(len=16)            :: mode
   integer                      :: nb, nft, seed, ios, nline, plen, i, j
   integer(c_int64_t)           :: nwarm, nmoves
   real(c_double)               :: taudiffmax
   real(c_double), allocatable  :: umat(:,:,:,:), ftau(:,:,:,:,:), muimp(:,:,:,:), screen(:,:,:,:,:)
   character(len=:), allocatable :: pstr
   character(len=600)           :: line

   call get_command_argument(1, dir)

   open(10, file=trim(dir)//'/run.txt', status='old', action='read')
   read(10, *) mode
   read(10, *) nb
   read(10, *) nft
   read(10, *) seed
   read(10, *) nwarm
   read(10, *) nmoves
   read(10, *) taudiffmax
   close(10)

   ! parameter string: the file's lines joined with newlines (Parameters.F90:281-320)
   pstr = ''
   open(11, file=trim(dir)//'/params.txt', status='old', action='read')
   nline = 0
   do
      read(11, '(A)', iostat=ios) line
      if (ios /= 0) exit
      pstr = pstr // trim(line) // char(10)
      nline = nline + 1
   end do
   close(11)
   plen = len(pstr)

   allocate(umat(2*nb, 2*nb, 2*nb, 2*nb))
   allocate(ftau(nb, 2, nb, 2, nft), screen(nb, 2, nb, 2, nft))
   allocate(muimp(nb, 2, nb, 2))
   screen = 0
   call read_array(trim(dir)//'/umatrix.bin', umat, size(umat))
   call read_array(trim(dir)//'/ftau.bin', ftau, size(ftau))
   call read_array(trim(dir)//'/muimp.bin', muimp, size(muimp))

   call init_paras(pstr)
   call init_rnd(seed)
   call init_solver(umat, ftau, muimp, screen, 1, nft, nb)
   if (taudiffmax > 0) call set_taudiffmax(taudiffmax)

   call write_model(trim(dir))

   select case (trim(mode))
   case ('log')
      call run_log(trim(dir), nmoves)
   case ('run')
      call ctqmc_warmup(nwarm)
      call ctqmc_measure(1, 1)
      call write_results(trim(dir))
   case ('eval')
      call run_eval(trim(dir))
   case ('calib')
      call run_calib(trim(dir), nwarm)
   case ('runcfg')
      ! a caller may start the (slow) model setup early and release the run
      ! once configs.bin is written: one line on standard input (end of input
      ! releases it as well)
      read(*, '(A)', iostat=ios) line
      call start_from_config(trim(dir))
      call ctqmc_warmup(nwarm)
      call ctqmc_measure(1, 1)
      call write_results(trim(dir))
      open(80, file=trim(dir)//'/timing.txt', action='write', status='replace')
      write(80, '(2ES25.17, I10)') time_warmup, time_sim, DTrace%NOper
      close(80)
   case default
      stop 'unknown mode'
   end select

contains

   subroutine read_array(fname, arr, n)
      character(*), intent(in)      :: fname
      integer, intent(in)           :: n
      real(c_double), intent(inout) :: arr(n)
      open(20, file=fname, access='stream', form='unformatted', status='old', action='read')
      read(20) arr
      close(20)
   end subroutine read_array

   ! superstate sizes, eigenvalues, Connect table, Egs, window and NSlide
   subroutine write_model(d)
      character(*), intent(in) :: d
      integer :: s, b, sp, ca
      open(30, file=d//'/model.txt', action='write', status='replace')
      write(30, *) DStates%NSStates, DStates%NStates, DStates%NStatesMax
      write(30, '(ES25.17)') DTrace%Egs
      write(30, *) DTrace%num_windows, NSlide
      write(30, '(ES25.17)') DTrace%taudiff_max
      write(30, *) DTrace%outer_sst
      do s = 0, DStates%NSStates - 1
         write(30, *) s, DStates%SubStates(s)%NStates, DStates%SubStates(s)%Offset
         write(30, *) DStates%SubStates(s)%States(:)
         write(30, '(1000ES25.17)') DStates%SubStates(s)%EVal(:)
         write(30, *) (((DStates%SubStates(s)%Connect(b, sp, ca), b = 1, nb), sp = 1, 2), ca = 1, 2)
      end do
      close(30)
   end subroutine write_model

   ! one record per move: kind, outcome, noper, outer_sst, window_position,
   ! trace log/sign, det log/sign, cfgsign
   subroutine run_log(d, n)
      character(*), intent(in)      :: d
      integer(c_int64_t), intent(in) :: n
      integer(c_int64_t) :: imove
      integer(c_int32_t) :: kind, outcome
      real(c_double)     :: rand_num, acc_before, qn_before
      real(c_double)     :: pshift, pglob

      call ctqmc_warmup(0_c_int64_t)     ! resets counters and AccPair as a real run does
      pshift = PercentageTauShiftMove
      pglob = PercentageGlobalMove
      open(40, file=d//'/moves.bin', access='stream', form='unformatted', status='replace', action='write')
      do imove = 1, n
         rand_num = grnd()
         if (rand_num < (1d0 - pshift - pglob)/2d0) then
            kind = 1
            acc_before = AccAdd(1); qn_before = real(AccQNAdd(1), c_double)
            if (grnd() >= Percentage4OperatorMove) then
               call StepAdd(grnd() < PercentageOuterMove, 1)
            else
               stop 'four-operator moves are not exercised'
            end if
            outcome = merge(1, 0, AccAdd(1) > acc_before)
            if (real(AccQNAdd(1), c_double) > qn_before) outcome = 2
         else if (rand_num <= 1d0 - pshift - pglob) then
            kind = 2
            acc_before = AccRem(1); qn_before = real(AccQNRem(1), c_double)
            if (grnd() >= Percentage4OperatorMove) then
               call StepRem(grnd() < PercentageOuterMove, 1)
            else
               stop 'four-operator moves are not exercised'
            end if
            outcome = merge(1, 0, AccRem(1) > acc_before)
            if (real(AccQNRem(1), c_double) > qn_before) outcome = 2
         else if (rand_num <= 1d0 - pshift) then
            kind = 3
            acc_before = AccGlob; qn_before = real(AccQNGlob, c_double)
            call StepGlob()
            outcome = merge(1, 0, AccGlob > acc_before)
            if (real(AccQNGlob, c_double) > qn_before) outcome = 2
         else
            kind = 4
            acc_before = AccShift
            call StepShiftTau()
            outcome = merge(1, 0, AccShift > acc_before)
         end if
         if (modulo(sum(TryAdd) + sum(TryRem), NSlide) == 0) call shift_window(DTrace)
         write(40) kind, outcome, int(DTrace%NOper, c_int32_t), int(DTrace%outer_sst, c_int32_t), &
                   int(DTrace%window_position, c_int32_t), &
                   DTrace%Trace%log, DTrace%Trace%sign, DTrace%Det%log, DTrace%Det%sign, DTrace%cfgsign
      end do
      close(40)
      call write_config(d//'/final_config.bin')
   end subroutine run_log

   ! configuration + inverse hybridisation blocks in flavour order (orbital fastest)
   subroutine write_config(fname)
      character(*), intent(in) :: fname
      integer :: n, osst, ostate, b, sp
      real(c_double), allocatable :: taus(:)
      integer, allocatable :: orbs(:), spins(:), cas(:), hyb(:)
      call get_mc_config_scalars(n, osst, ostate)
      allocate(taus(n), orbs(n), spins(n), cas(n), hyb(n))
      call get_mc_config(n, taus, orbs, spins, cas, hyb, osst, ostate)
      open(50, file=fname, access='stream', form='unformatted', status='replace', action='write')
      write(50) int(n, c_int32_t), int(osst, c_int32_t)
      write(50) taus
      write(50) int(orbs, c_int32_t), int(spins, c_int32_t), int(cas, c_int32_t)
      write(50) DTrace%Trace%log, DTrace%Trace%sign, DTrace%Det%log, DTrace%Det%sign, DTrace%cfgsign
      do sp = 1, 2
         do b = 1, nb
            write(50) int(size(DTrace%MInv(b, sp)%Mat, 1), c_int32_t)
            write(50) DTrace%MInv(b, sp)%Mat
         end do
      end do
      close(50)
   end subroutine write_config

   ! configs.bin: int32 count, then per config: int32 n, int32 outer_sst,
   ! n doubles tau, 3n int32 (orb, spin, ca)
   subroutine run_eval(d)
      character(*), intent(in) :: d
      integer(c_int32_t) :: ncfg, n, osst, icfg
      real(c_double), allocatable :: taus(:)
      integer(c_int32_t), allocatable :: o(:), s(:), c(:)
      integer, allocatable :: hyb(:)
      integer :: b, sp
      open(60, file=d//'/configs.bin', access='stream', form='unformatted', status='old', action='read')
      open(61, file=d//'/eval.bin', access='stream', form='unformatted', status='replace', action='write')
      read(60) ncfg
      do icfg = 1, ncfg
         read(60) n, osst
         allocate(taus(n), o(n), s(n), c(n), hyb(n))
         read(60) taus
         read(60) o, s, c
         hyb = -1
         call set_mc_config(int(n), taus, int(o), int(s), int(c), hyb, int(osst), 1)
         write(61) DTrace%Trace%log, DTrace%Trace%sign, DTrace%Det%log, DTrace%Det%sign, DTrace%cfgsign
         do sp = 1, 2
            do b = 1, nb
               write(61) int(size(DTrace%MInv(b, sp)%Mat, 1), c_int32_t)
               write(61) DTrace%MInv(b, sp)%Mat
            end do
         end do
         deallocate(taus, o, s, c, hyb)
      end do
      close(60)
      close(61)
   end subroutine run_eval

   ! the calibration of cthyb_solver.py:899-944 for one rank (phase 1 list,
   ! TaudiffMax from its 99.5 % point, phase 2 histogram; the fit of phase 2 is
   ! the driver's and is not needed here), then warm-up and measurement; writes
   ! calib.bin: phase-1 list, taudiffmax, phase-2 AccPair, then after the
   ! measurement AccPair, the outer-state histograms and the step arrays
   subroutine run_calib(d, nw)
      character(*), intent(in) :: d
      integer(c_int64_t), intent(in) :: nw
      real(c_double), allocatable :: lst(:), lst0(:)
      real(c_double) :: tdm
      integer :: n1
      n1 = 1000
      call init_counters()
      call ctqmc_calibrate(.true., n1, 100, 1000, 100000)
      lst0 = AccPairTau
      lst = lst0
      call sort_real(lst)
      tdm = lst(int(0.995d0*n1) + 1)
      call set_taudiffmax(tdm)
      call init_counters()
      call ctqmc_calibrate(.false., n1, 100, 1000, 100000)
      open(90, file=d//'/calib.bin', access='stream', form='unformatted', status='replace', action='write')
      write(90) int(size(lst0), c_int32_t)
      write(90) lst0
      write(90) tdm
      write(90) int(size(AccPair, 3), c_int32_t)
      write(90) real(AccPair, c_double)
      call ctqmc_warmup(nw)
      call ctqmc_measure(1, 1)
      write(90) int(size(AccPair, 3), c_int32_t)
      write(90) real(AccPair, c_double)
      write(90) OuterStateHisto, SignHistoSuperstates, SignHistoStates, TraceContribSuperstates, &
                TraceContribStates
      write(90) Gtau_mean_step, Gtau_mid_step, sign_step
      write(90) rho1, rho2
      close(90)
      call write_results(d)
   end subroutine run_calib

   subroutine sort_real(a)
      real(c_double), intent(inout) :: a(:)
      integer :: i, j
      real(c_double) :: x
      do i = 2, size(a)
         x = a(i)
         j = i - 1
         do while (j >= 1)
            if (a(j) <= x) exit
            a(j + 1) = a(j)
            j = j - 1
         end do
         a(j + 1) = x
      end do
   end subroutine sort_real

   subroutine start_from_config(d)
      character(*), intent(in) :: d
      integer(c_int32_t) :: ncfg, n, osst
      real(c_double), allocatable :: taus(:)
      integer(c_int32_t), allocatable :: o(:), s(:), c(:)
      integer, allocatable :: hyb(:)
      open(60, file=d//'/configs.bin', access='stream', form='unformatted', status='old', action='read')
      read(60) ncfg
      read(60) n, osst
      allocate(taus(n), o(n), s(n), c(n), hyb(n))
      read(60) taus
      read(60) o, s, c
      close(60)
      hyb = -1
      call set_mc_config(int(n), taus, int(o), int(s), int(c), hyb, int(osst), 1)
   end subroutine start_from_config

   subroutine write_results(d)
      character(*), intent(in) :: d
      open(70, file=d//'/results.bin', access='stream', form='unformatted', status='replace', action='write')
      write(70) mean_sign
      write(70) gtau
      write(70) gleg
      write(70) occ
      write(70) densitymatrix
      write(70) histo
      write(70) AccAdd(1), AccRem(1), AccGlob, AccShift
      write(70) OuterSuperstateHisto
      close(70)
   end subroutine write_results

end program ref_driver
