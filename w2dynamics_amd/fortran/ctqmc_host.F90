!===============================================================================
!> Fortran host of the MI355X CT-HYB sampling core.
!!
!! Module MCTQMC keeps the entry points, argument meanings and result arrays
!! that the w2dyn Python driver uses from the reference solver module
!! (w2dynamics src/ctqmc_fortran/CTQMC.F90; f2py list in
!! src/ctqmc_fortran/CMakeLists.txt:61-72, calls in
!! w2dyn/dmft/cthyb_solver.py:240-342 and reads in :949-1030).  Every call goes
!! through the C-ABI of libw2dyn_ctqmc.so (include/w2dyn_ctqmc.h) by
!! iso_c_binding; no sampling happens here.
!!
!! Differences from the reference module, all on purpose:
!!  - one handle runs n_chains independent Markov chains on one GPU (chain c
!!    seeded with Nseed + c, like MPI ranks NSeed + rank); the result arrays are
!!    the normalisation of CTQMC.F90:8663-8771 applied to the sums over all
!!    chains (the sign-ratio estimate sum(G sign) / sum(sign), finite where a
!!    single chain's sign sum vanishes).  n_chains = 1 (default) is the
!!    reference's single-rank run move for move;
!!  - errors are returned by the library as status codes; this layer prints
!!    the code and sets `aborted` = 2 instead of stopping the process; chains
!!    that stop on an error while others complete are left out of the results
!!    (a warning, not an abort);
!!  - only the Z-space sector (iSector = 1) with diagonal hybridisation is
!!    available; other sectors set aborted = 2.
!===============================================================================
module MCTQMC
   use iso_c_binding
   implicit none

   integer, parameter :: KINDR = kind(0.d0)

   !> batching controls (no counterpart in the reference module)
   integer :: n_chains = 1
   integer :: chain_capacity = 256
   integer :: device_id = 0
   integer :: config_chain = 0

   !> reference module variables used by the driver
   integer :: SimID = -1
   integer :: aborted = 0
   logical :: fancy_prog = .false.
   real(KINDR) :: time_warmup = 0, time_sim = 0, time_sim_steps = 0
   real(KINDR) :: mean_sign = 0
   real(KINDR), allocatable :: Gtau(:,:,:)           ! (NBands,2,Ntau)
   real(KINDR), allocatable :: GLeg(:,:,:)           ! (NBands,2,NLegMax)
   real(KINDR), allocatable :: occ(:,:,:,:)          ! (NBands,2,NBands,2)
   real(KINDR), allocatable :: DensityMatrix(:,:)    ! (NStates,NStates), occupation basis
   real(KINDR), allocatable :: Histo(:,:,:)          ! (NBands,2,0:MaxHisto)
   real(KINDR), allocatable :: OuterSuperstateHisto(:)
   real(KINDR), allocatable :: EigenstatesEnergies(:)
   integer, allocatable :: StatesSuperstates(:)      ! superstate of each eigen index
   integer, allocatable :: OccbasisMapping(:)        ! occupation-basis state of each index
   real(KINDR) :: AccAdd(1) = 0, AccRem(1) = 0, AccGlob = 0, AccShift = 0
   integer(c_int64_t) :: TryAdd(1) = 0, TryRem(1) = 0, TryGlob = 0, TryShift = 0
   !> four-operator and flavour-change moves are not taken by this core: zero
   real(KINDR) :: AccAdd4 = 0, AccRem4 = 0, AccFlavc = 0
   !> accepted-pair statistics (CTQMC.F90:66-70): AccPair is the mean over
   !! chains of the counts per (orbital, spin, distance bin); AccPairTau the
   !! mean over chains of each chain's sorted phase-1 distance list, so that
   !! its 99.5 % point is the mean of the chains' 99.5 % points, the value the
   !! driver forms over ranks (cthyb_solver.py:911-915)
   real(KINDR), allocatable :: AccPair(:,:,:)
   real(KINDR), allocatable :: AccPairTau(:)
   !> MeasOuterHistograms outputs (CTQMC.F90:3285-3339, normalised as 8757-8764)
   real(KINDR), allocatable :: OuterStateHisto(:), SignHistoSuperstates(:), SignHistoStates(:)
   real(KINDR), allocatable :: TraceContribSuperstates(:), TraceContribStates(:)
   !> reduced density matrices (CTQMC.F90:2936-3060), flavour index 2*(b-1)+s
   real(KINDR), allocatable :: rho1(:,:), rho2(:,:,:,:)
   !> off-diagonal G(tau) / Legendre G and the Lanczos histograms: paths this
   !! core does not take, zero as in the reference for diagonal hybridisation
   real(KINDR), allocatable :: Gtau_full(:,:,:,:,:), GLeg_full(:,:,:,:,:)
   real(KINDR) :: lhisto(100) = 0, rhisto(100) = 0
   !> step-resolved arrays (Gtau_mean_step, Gtau_mid_step, sign_step;
   !! CTQMC.F90:2394-2402, 8741-8742), mean over chains; allocated if enabled
   real(KINDR), allocatable :: Gtau_mean_step(:,:,:), Gtau_mid_step(:,:,:), sign_step(:)
   !> AccumGtau = 2: binning variances of the raw G(tau) per level
   !! (gtau_blocks, CTQMC.F90:8733-8739), mean over chains
   real(KINDR), allocatable :: gtau_blocks(:,:,:,:)
   !> MeasGiw: G(i w_n), n = -Niw..Niw-1 (Giw, CTQMC.F90:573, normalised as
   !! 8816) from the sums over all chains; allocated if enabled
   complex(KINDR), allocatable :: Giw(:,:,:)        ! (NBands,2,-Niw:Niw-1)
   real(KINDR) :: time_calibration = 0

   type(c_ptr), private, save :: handle = c_null_ptr
   character(len=:), allocatable, private, save :: paramstring
   integer, private, save :: seed = 0
   integer, private, save :: nb = 0, ntau = 0, nleg = 0, nsst = 0, nstates = 0, maxhisto = 0
   real(KINDR), private, save :: beta = 0
   integer(c_int), allocatable, private, save :: model_sizes(:), model_offs(:)
   real(KINDR), allocatable, private, save :: model_vec(:)

   ! internal helpers and the C-ABI bindings are not part of the module's
   ! interface (an f2py wrapper of the module sees the reference's names only)
   private :: failed, param_value, load_model, collect_accpair, sort_ascending, collect_results, &
              collect_histograms, apply_ops, reduced_density_matrices, collect_steps_and_blocks, &
              reset_results, collect_giw
   private :: w2d_ctqmc_aborted, w2d_ctqmc_calibrate, w2d_ctqmc_create, w2d_ctqmc_destroy, &
              w2d_ctqmc_dims, w2d_ctqmc_get_accumulators, w2d_ctqmc_get_config, &
              w2d_ctqmc_get_group_histograms, w2d_ctqmc_get_group_sums, w2d_ctqmc_get_gtau_blocks, &
              w2d_ctqmc_get_giw, &
              w2d_ctqmc_get_model, w2d_ctqmc_get_pair_stats, w2d_ctqmc_get_steps, &
              w2d_ctqmc_init_counters, w2d_ctqmc_measure, w2d_ctqmc_set_config, &
              w2d_ctqmc_set_taudiffmax, w2d_ctqmc_trap_signals, w2d_ctqmc_warmup

   interface
      integer(c_int) function w2d_ctqmc_create(params, nbands, nftau, u, ftau, muimp, nchains, &
                                               seed0, capacity, device, out) &
            bind(C, name="w2d_ctqmc_create")
         import :: c_int, c_char, c_double, c_int32_t, c_ptr
         character(kind=c_char), dimension(*), intent(in) :: params
         integer(c_int), value :: nbands, nftau
         real(c_double), dimension(*), intent(in) :: u, ftau, muimp
         integer(c_int), value :: nchains
         integer(c_int32_t), value :: seed0
         integer(c_int), value :: capacity, device
         type(c_ptr), intent(out) :: out
      end function
      integer(c_int) function w2d_ctqmc_set_taudiffmax(h, x) bind(C, name="w2d_ctqmc_set_taudiffmax")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), value :: x
      end function
      integer(c_int) function w2d_ctqmc_warmup(h, n) bind(C, name="w2d_ctqmc_warmup")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: h
         integer(c_int64_t), value :: n
      end function
      integer(c_int) function w2d_ctqmc_measure(h, n) bind(C, name="w2d_ctqmc_measure")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: h
         integer(c_int64_t), value :: n
      end function
      integer(c_int) function w2d_ctqmc_dims(h, nbands, nchains, ntau, nlegmax, nsst, nstates, &
                                             maxhisto) bind(C, name="w2d_ctqmc_dims")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), intent(out) :: nbands, nchains, ntau, nlegmax, nsst, nstates, maxhisto
      end function
      integer(c_int) function w2d_ctqmc_get_accumulators(h, gtau, gleg, dm, histo, outer, scalars) &
            bind(C, name="w2d_ctqmc_get_accumulators")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), dimension(*), intent(out) :: gtau, gleg, dm, histo, outer, scalars
      end function
      integer(c_int) function w2d_ctqmc_get_model(h, sizes, offsets, evals, states, evecs, egs) &
            bind(C, name="w2d_ctqmc_get_model")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         type(c_ptr), value :: sizes, offsets, evals, states, evecs, egs   ! NULL = not requested
      end function
      integer(c_int) function w2d_ctqmc_get_config(h, chain, cap, n, tau, orb, spin, ca, outer_sst, &
                                                   weights) bind(C, name="w2d_ctqmc_get_config")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: chain, cap
         integer(c_int), intent(out) :: n, outer_sst
         real(c_double), dimension(*), intent(out) :: tau, weights
         integer(c_int), dimension(*), intent(out) :: orb, spin, ca
      end function
      integer(c_int) function w2d_ctqmc_set_config(h, chain, n, tau, orb, spin, ca, outer_sst, &
                                                   weights) bind(C, name="w2d_ctqmc_set_config")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: chain, n, outer_sst
         real(c_double), dimension(*), intent(in) :: tau
         integer(c_int), dimension(*), intent(in) :: orb, spin, ca
         real(c_double), dimension(*), intent(out) :: weights
      end function
      integer(c_int) function w2d_ctqmc_get_group_sums(h, ngroups, gtau, gleg, dmb, histo, outer, &
                                                       scalars, nused) bind(C, name="w2d_ctqmc_get_group_sums")
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: h
         integer(c_int), value :: ngroups
         real(c_double) :: gtau(*), gleg(*), dmb(*), histo(*), outer(*), scalars(*)
         integer(c_int32_t) :: nused(*)
      end function
      integer(c_int) function w2d_ctqmc_aborted(h, ab) bind(C, name="w2d_ctqmc_aborted")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int) :: ab
      end function
      integer(c_int) function w2d_ctqmc_trap_signals() bind(C, name="w2d_ctqmc_trap_signals")
         import :: c_int
      end function
      integer(c_int) function w2d_ctqmc_init_counters(h) bind(C, name="w2d_ctqmc_init_counters")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function w2d_ctqmc_calibrate(h, list, p1, p2grid, p2pairs, nprog) &
            bind(C, name="w2d_ctqmc_calibrate")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: list, p1, p2grid, p2pairs, nprog
      end function
      integer(c_int) function w2d_ctqmc_get_pair_stats(h, chain, apt, napt, ap, nap) &
            bind(C, name="w2d_ctqmc_get_pair_stats")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: chain
         type(c_ptr), value :: apt, ap     ! c_null_ptr: sizes only
         integer(c_int) :: napt, nap
      end function
      integer(c_int) function w2d_ctqmc_get_group_histograms(h, ng, ost, ssst, sst, csst, cst, ap) &
            bind(C, name="w2d_ctqmc_get_group_histograms")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: ng
         real(c_double) :: ost(*), ssst(*), sst(*), csst(*), cst(*), ap(*)
      end function
      integer(c_int) function w2d_ctqmc_get_steps(h, chain, nm, ms, ds, ss) bind(C, name="w2d_ctqmc_get_steps")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: chain
         integer(c_int) :: nm
         type(c_ptr), value :: ms, ds, ss  ! c_null_ptr: size only
      end function
      integer(c_int) function w2d_ctqmc_get_gtau_blocks(h, chain, nl, blocks) bind(C, name="w2d_ctqmc_get_gtau_blocks")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: chain
         integer(c_int) :: nl
         type(c_ptr), value :: blocks      ! c_null_ptr: size only
      end function
      integer(c_int) function w2d_ctqmc_get_giw(h, ngroups, niw, out) bind(C, name="w2d_ctqmc_get_giw")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: ngroups
         integer(c_int) :: niw
         type(c_ptr), value :: out         ! c_null_ptr: size only
      end function
      integer(c_int) function w2d_ctqmc_destroy(h) bind(C, name="w2d_ctqmc_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
   end interface

contains

   !> status check: print and flag instead of stopping (the reference stops);
   !! W2D_ERR_PARTIAL (some chains stopped on an error, the others completed)
   !! is a warning: the stopped chains are left out of the results
   logical function failed(st, where)
      integer(c_int), intent(in) :: st
      character(*), intent(in) :: where
      failed = st /= 0 .and. st /= 7
      if (st == 7) write (0, '(A,A,A)') "MCTQMC: ", where, &
         ": some chains stopped on an error and are left out of the results"
      if (failed) then
         write (0, '(A,A,A,I0)') "MCTQMC: ", where, " returned status ", st
         aborted = 2
      end if
   end function failed

   !> value of `key` in the parameter string (first line whose key starts with it)
   function param_value(key) result(val)
      character(*), intent(in) :: key
      character(len=:), allocatable :: val
      integer :: i0, i1, ieq, ih
      character(len=:), allocatable :: line
      val = ""
      i0 = 1
      do while (i0 <= len(paramstring))
         i1 = index(paramstring(i0:), new_line('a'))
         if (i1 == 0) then
            line = paramstring(i0:)
            i0 = len(paramstring) + 1
         else
            line = paramstring(i0:i0 + i1 - 2)
            i0 = i0 + i1
         end if
         ieq = index(line, "=")
         if (ieq == 0) cycle
         if (index(adjustl(line(:ieq - 1)), key) /= 1) cycle
         ih = index(line(ieq + 1:), "#")
         if (ih == 0) then
            val = trim(adjustl(line(ieq + 1:)))
         else
            val = trim(adjustl(line(ieq + 1:ieq + ih - 1)))
         end if
         return
      end do
   end function param_value

   !> init_paras (CTQMC.F90:9412): the "key = value" parameter string
   subroutine init_paras(parastr)
      character(*), intent(in) :: parastr
      paramstring = parastr
   end subroutine init_paras

   !> dest_paras (CTQMC.F90:9420)
   subroutine dest_paras()
      if (allocated(paramstring)) deallocate (paramstring)
   end subroutine dest_paras

   !> init_rnd (CTQMC.F90:9404): chain c is seeded with Nseed + c
   subroutine init_rnd(Nseed)
      integer, intent(in) :: Nseed
      seed = Nseed
   end subroutine init_rnd

   !> init_solver (CTQMC.F90:7894-8085): local model, hybridisation and chains.
   !! screening_function and iter_no are accepted for the same call shape;
   !! retarded interactions are not supported by the library.
   subroutine init_solver(u_matrix_in, Ftau_full, muimp_full, screening_function, iter_no, Nftau, NBands)
      integer, intent(in) :: iter_no, Nftau, NBands
      real(KINDR), intent(in) :: u_matrix_in(2*NBands, 2*NBands, 2*NBands, 2*NBands)
      real(KINDR), intent(in) :: Ftau_full(NBands, 2, NBands, 2, Nftau)
      real(KINDR), intent(in) :: muimp_full(NBands, 2, NBands, 2)
      real(KINDR), intent(in) :: screening_function(NBands, 2, NBands, 2, Nftau)
      integer(c_int) :: st, nbr, ncr
      character(len=:), allocatable :: bs
      if (.not. allocated(paramstring)) then
         write (0, '(A)') "MCTQMC: init_paras must precede init_solver"
         aborted = 2
         return
      end if
      if (c_associated(handle)) call dest_ctqmc()
      aborted = 0
      st = w2d_ctqmc_create(paramstring//c_null_char, int(NBands, c_int), int(Nftau, c_int), &
                            u_matrix_in, Ftau_full, muimp_full, int(n_chains, c_int), &
                            int(seed, c_int32_t), int(chain_capacity, c_int), int(device_id, c_int), handle)
      if (failed(st, "w2d_ctqmc_create")) then
         handle = c_null_ptr
         return
      end if
      st = w2d_ctqmc_dims(handle, nbr, ncr, ntau, nleg, nsst, nstates, maxhisto)
      if (failed(st, "w2d_ctqmc_dims")) return
      nb = nbr
      bs = param_value("beta")
      read (bs, *) beta
      call load_model()
      ! register_trap for SIGTERM, SIGINT, SIGUSR1 (CTQMC.F90:8076-8078)
      if (failed(w2d_ctqmc_trap_signals(), "w2d_ctqmc_trap_signals")) return
      if (.false.) print *, iter_no, screening_function(1, 1, 1, 1, 1)
   end subroutine init_solver

   subroutine load_model()
      integer(c_int), allocatable, target :: sizes(:), offs(:), states(:)
      real(KINDR), allocatable, target :: ev(:), vec(:)
      real(KINDR), target :: egs(1)
      integer(c_int) :: st
      integer :: s, i
      allocate (sizes(nsst), offs(nsst), states(nstates), ev(nstates))
      ! block sizes first, then the eigenvector blocks (column-major, concatenated)
      st = w2d_ctqmc_get_model(handle, c_loc(sizes), c_loc(offs), c_loc(ev), c_loc(states), &
                               c_null_ptr, c_loc(egs))
      if (failed(st, "w2d_ctqmc_get_model")) return
      allocate (vec(max(1_8, sum(int(sizes, 8)**2))))
      st = w2d_ctqmc_get_model(handle, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(vec), &
                               c_null_ptr)
      if (failed(st, "w2d_ctqmc_get_model")) return
      EigenstatesEnergies = ev
      OccbasisMapping = states
      if (allocated(StatesSuperstates)) deallocate (StatesSuperstates)
      allocate (StatesSuperstates(nstates))
      do s = 1, nsst
         do i = 1, sizes(s)
            StatesSuperstates(offs(s) + i) = s - 1
         end do
      end do
      model_sizes = sizes
      model_offs = offs
      model_vec = vec
   end subroutine load_model

   !> set_taudiffmax (CTQMC.F90:8180)
   subroutine set_taudiffmax(taudiffmax)
      real(KINDR), intent(in) :: taudiffmax
      if (failed(w2d_ctqmc_set_taudiffmax(handle, taudiffmax), "w2d_ctqmc_set_taudiffmax")) return
   end subroutine set_taudiffmax

   !> init_counters (CTQMC.F90:847-899): move counters, AccPair, calibration list index
   subroutine init_counters()
      if (failed(w2d_ctqmc_init_counters(handle), "w2d_ctqmc_init_counters")) return
   end subroutine init_counters

   !> ctqmc_calibrate (CTQMC.F90:8088-8178) on every chain; fills AccPairTau
   !! (list) or AccPair (phase 2) as described at their declarations
   subroutine ctqmc_calibrate(list, phase1pairnum, phase2taugrid, phase2pairnum, nprogress)
      logical, intent(in) :: list
      integer, intent(in) :: phase1pairnum, phase2taugrid, phase2pairnum, nprogress
      integer(8) :: t0, t1, rate
      integer :: c
      integer(c_int) :: napt, nap, lflag
      real(KINDR), allocatable, target :: lst(:)
      call system_clock(t0, rate)
      lflag = merge(1, 0, list)
      if (failed(w2d_ctqmc_calibrate(handle, lflag, int(phase1pairnum, c_int), int(phase2taugrid, c_int), &
                                     int(phase2pairnum, c_int), int(nprogress, c_int)), &
                 "w2d_ctqmc_calibrate")) return
      call system_clock(t1)
      time_calibration = time_calibration + real(t1 - t0, KINDR)/real(rate, KINDR)
      if (list) then
         if (allocated(AccPairTau)) deallocate (AccPairTau)
         allocate (AccPairTau(phase1pairnum), lst(phase1pairnum))
         AccPairTau = 0
         do c = 0, n_chains - 1
            lst = 0
            if (failed(w2d_ctqmc_get_pair_stats(handle, int(c, c_int), c_null_ptr, napt, c_null_ptr, nap), &
                       "w2d_ctqmc_get_pair_stats")) return
            if (napt > phase1pairnum) then
               write (*, '(A)') 'MCTQMC: accepted-pair list longer than phase1pairnum'
               aborted = 2
               return
            end if
            if (failed(w2d_ctqmc_get_pair_stats(handle, int(c, c_int), c_loc(lst), napt, c_null_ptr, nap), &
                       "w2d_ctqmc_get_pair_stats")) return
            call sort_ascending(lst(1:napt))
            AccPairTau = AccPairTau + lst/n_chains
         end do
      else
         if (allocated(AccPairTau)) deallocate (AccPairTau)
         call collect_accpair()
      end if
   end subroutine ctqmc_calibrate

   !> AccPair: mean over chains of the counts
   subroutine collect_accpair()
      integer :: c
      integer(c_int) :: napt, nap
      real(KINDR), allocatable, target :: ap(:,:)
      if (failed(w2d_ctqmc_get_pair_stats(handle, 0_c_int, c_null_ptr, napt, c_null_ptr, nap), &
                 "w2d_ctqmc_get_pair_stats")) return
      if (allocated(AccPair)) deallocate (AccPair)
      allocate (AccPair(nb, 2, nap), ap(nap, 2*nb))
      AccPair = 0
      do c = 0, n_chains - 1
         if (failed(w2d_ctqmc_get_pair_stats(handle, int(c, c_int), c_null_ptr, napt, c_loc(ap), nap), &
                    "w2d_ctqmc_get_pair_stats")) return
         AccPair(:, 1, :) = AccPair(:, 1, :) + transpose(ap(:, 1:nb))/n_chains
         AccPair(:, 2, :) = AccPair(:, 2, :) + transpose(ap(:, nb + 1:2*nb))/n_chains
      end do
   end subroutine collect_accpair

   subroutine sort_ascending(a)
      real(KINDR), intent(inout) :: a(:)
      integer :: i, j
      real(KINDR) :: x
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
   end subroutine sort_ascending

   !> ctqmc_warmup (CTQMC.F90:8188): Nwarmups steps of every chain
   subroutine ctqmc_warmup(Nwarmups)
      integer(c_int64_t), intent(in) :: Nwarmups
      integer(8) :: t0, t1, rate
      call system_clock(t0, rate)
      if (failed(w2d_ctqmc_warmup(handle, Nwarmups), "w2d_ctqmc_warmup")) return
      call system_clock(t1)
      time_warmup = real(t1 - t0, KINDR)/real(rate, KINDR)
   end subroutine ctqmc_warmup

   !> ctqmc_measure (CTQMC.F90:8355-8771), Z-space sector: Nmeas x (NCorr steps +
   !! measurements) per chain, then the normalisation of CTQMC.F90:8663-8771
   !! per chain and the mean over chains into the module arrays.
   subroutine ctqmc_measure(iSector, iComponent)
      integer, intent(in) :: iSector, iComponent
      integer(c_int64_t) :: nmeas
      integer(8) :: t0, t1, rate
      integer(c_int) :: ab
      character(len=:), allocatable :: v
      if (iSector /= 1) then
         write (0, '(A,I0,A,I0)') "MCTQMC: only the Z-space sector is available, got sector ", &
            iSector, " component ", iComponent
         aborted = 2
         return
      end if
      v = param_value("Nmeas")
      read (v, *) nmeas
      call system_clock(t0, rate)
      if (failed(w2d_ctqmc_measure(handle, nmeas), "w2d_ctqmc_measure")) return
      call system_clock(t1)
      time_sim = real(t1 - t0, KINDR)/real(rate, KINDR)
      time_sim_steps = time_sim
      ! a trapped signal ends the sampling early (CTQMC.F90:8648-8658)
      if (failed(w2d_ctqmc_aborted(handle, ab), "w2d_ctqmc_aborted")) return
      aborted = ab
      if (ab == 2) return
      call collect_results()
   end subroutine ctqmc_measure

   subroutine collect_results()
      ! sums over all chains that completed (one group), then the normalisation
      ! of CTQMC.F90:8663-8771 once: a ratio of sums, finite where one chain's
      ! sign sum vanishes
      real(KINDR), allocatable :: g(:,:), gl(:,:), dmb(:), hi(:,:), oh(:), v(:,:), blk(:,:), diag(:)
      real(KINDR) :: sc(14), cnt, msign, zs, lfac
      integer(c_int32_t) :: nu(1)
      integer :: b, s, f, l, k, o, n, b2, s2, voff, boff
      integer(c_int) :: st
      allocate (g(ntau, 2*nb), gl(nleg, 2*nb), dmb(max(1, sum(model_sizes**2))), hi(maxhisto + 1, 2*nb))
      allocate (oh(nsst))
      st = w2d_ctqmc_get_group_sums(handle, 1_c_int, g, gl, dmb, hi, oh, sc, nu)
      if (failed(st, "w2d_ctqmc_get_group_sums")) return
      call reset_results()
      allocate (diag(nstates))
      cnt = sc(2)
      msign = sc(1)/cnt
      zs = -1.0_KINDR/(cnt*msign)
      mean_sign = msign
      do s = 1, 2
         do b = 1, nb
            f = b + nb*(s - 1)
            Gtau(b, s, :) = -(g(:, f)/cnt)/(-msign)*ntau/beta**2
            do l = 1, nleg
               lfac = sqrt(2.0_KINDR*(l - 1) + 1.0_KINDR)/beta
               GLeg(b, s, l) = -gl(l, f)*lfac*zs
            end do
            Histo(b, s, :) = hi(:, f)/cnt
         end do
      end do
      OuterSuperstateHisto = oh/cnt
      ! density matrix: eigenbasis blocks -> occupation basis (V rho V^T)
      voff = 0
      boff = 0
      do k = 1, nsst
         o = model_offs(k)
         n = model_sizes(k)
         v = reshape(model_vec(voff + 1:voff + n*n), [n, n])
         blk = reshape(dmb(boff + 1:boff + n*n), [n, n])*zs
         DensityMatrix(o + 1:o + n, o + 1:o + n) = matmul(v, matmul(blk, transpose(v)))
         voff = voff + n*n
         boff = boff + n*n
      end do
      AccAdd(1) = sc(3)/max(sc(4), 1.0_KINDR)
      AccRem(1) = sc(5)/max(sc(6), 1.0_KINDR)
      AccGlob = sc(7)/max(sc(8), 1.0_KINDR)
      AccShift = sc(9)/max(sc(10), 1.0_KINDR)
      TryAdd(1) = int(sc(4), c_int64_t)
      TryRem(1) = int(sc(6), c_int64_t)
      TryGlob = int(sc(8), c_int64_t)
      TryShift = int(sc(10), c_int64_t)
      call collect_histograms(cnt, oh)
      call collect_accpair()
      call reduced_density_matrices()
      call collect_steps_and_blocks()
      call collect_giw(cnt, msign)
      ! occ(b1,s1,b2,s2) = sum_k rho_kk n_{b1 s1}(k) n_{b2 s2}(k), occupation basis
      do k = 1, nstates
         diag(k) = DensityMatrix(k, k)
      end do
      do s2 = 1, 2
         do b2 = 1, nb
            do s = 1, 2
               do b = 1, nb
                  do k = 1, nstates
                     if (btest(OccbasisMapping(k), b - 1 + nb*(s - 1)) .and. &
                         btest(OccbasisMapping(k), b2 - 1 + nb*(s2 - 1))) &
                        occ(b, s, b2, s2) = occ(b, s, b2, s2) + diag(k)
                  end do
               end do
            end do
         end do
      end do
   end subroutine collect_results

   !> MeasOuterHistograms over all chains, normalised as CTQMC.F90:8757-8764
   subroutine collect_histograms(cnt, oh_raw)
      real(KINDR), intent(in) :: cnt, oh_raw(:)
      real(KINDR), allocatable :: ost(:), ssst(:), sst(:), csst(:), cst(:), ap(:)
      allocate (ost(nstates), ssst(nsst), sst(nstates), csst(nsst), cst(nstates))
      allocate (ap(2*nb*max(ntau, 1024)))
      if (failed(w2d_ctqmc_get_group_histograms(handle, 1_c_int, ost, ssst, sst, csst, cst, ap), &
                 "w2d_ctqmc_get_group_histograms")) return
      SignHistoSuperstates = ssst/oh_raw
      SignHistoStates = sst/ost
      OuterStateHisto = ost/cnt
      TraceContribSuperstates = csst/cnt
      TraceContribStates = cst/cnt
   end subroutine collect_histograms

   !> product of fermionic operators (applied right to left; flavour 2*(b-1)+s,
   !! dagger) on an occupation-basis state: bit (b-1) + NBands*(s-1), sign
   !! from the occupied lower bits (Operator.F90:160-172)
   subroutine apply_ops(nop, fl, dag, st, sg, ok)
      integer, intent(in) :: nop, fl(nop), dag(nop)
      integer, intent(inout) :: st
      real(KINDR), intent(out) :: sg
      logical, intent(out) :: ok
      integer :: i, bt
      sg = 1
      ok = .true.
      do i = nop, 1, -1
         bt = (fl(i) - 1)/2 + nb*mod(fl(i) - 1, 2)
         if (merge(1, 0, btest(st, bt)) == dag(i)) then
            ok = .false.
            return
         end if
         if (mod(popcnt(iand(st, ishft(1, bt) - 1)), 2) == 1) sg = -sg
         st = ieor(st, ishft(1, bt))
      end do
   end subroutine apply_ops

   !> rho1(f2,f1) = <c+_f2 c_f1>, rho2(f4,f3,f2,f1) = <c+_f4 c+_f3 c_f2 c_f1>
   !! (Meas_rho1/rho2_from_Densitymatrix, CTQMC.F90:2936-3060) from the
   !! occupation-basis density matrix: Tr[rho O] = sum_k rho(k, O(k)) sign_k
   subroutine reduced_density_matrices()
      integer, allocatable :: idx(:)
      integer :: nf, k, st, f1, f2, f3, f4
      real(KINDR) :: sg
      logical :: ok
      nf = 2*nb
      allocate (idx(0:2**nf - 1))
      idx = 0
      do k = 1, nstates
         idx(OccbasisMapping(k)) = k
      end do
      if (allocated(rho1)) deallocate (rho1, rho2)
      allocate (rho1(nf, nf), rho2(nf, nf, nf, nf))
      rho1 = 0
      rho2 = 0
      do f1 = 1, nf
         do f2 = 1, nf
            do k = 1, nstates
               st = OccbasisMapping(k)
               call apply_ops(2, [f2, f1], [1, 0], st, sg, ok)
               if (ok) rho1(f2, f1) = rho1(f2, f1) + DensityMatrix(k, idx(st))*sg
            end do
            do f3 = 1, nf
               do f4 = 1, nf
                  do k = 1, nstates
                     st = OccbasisMapping(k)
                     call apply_ops(4, [f4, f3, f2, f1], [1, 1, 0, 0], st, sg, ok)
                     if (ok) rho2(f4, f3, f2, f1) = rho2(f4, f3, f2, f1) + DensityMatrix(k, idx(st))*sg
                  end do
               end do
            end do
         end do
      end do
   end subroutine reduced_density_matrices

   !> step-resolved arrays and binning blocks (mean over chains)
   subroutine collect_steps_and_blocks()
      integer(c_int) :: nm, nl
      integer :: c, b, s, f
      real(KINDR), allocatable, target :: ms(:,:), ds(:,:), ss(:), bl(:,:,:)
      if (allocated(Gtau_mean_step)) deallocate (Gtau_mean_step)
      if (allocated(Gtau_mid_step)) deallocate (Gtau_mid_step)
      if (allocated(sign_step)) deallocate (sign_step)
      if (allocated(gtau_blocks)) deallocate (gtau_blocks)
      if (failed(w2d_ctqmc_get_steps(handle, 0_c_int, nm, c_null_ptr, c_null_ptr, c_null_ptr), &
                 "w2d_ctqmc_get_steps")) return
      if (nm > 0) then
         allocate (ms(2*nb, nm), ds(2*nb, nm), ss(nm))
         allocate (Gtau_mean_step(nb, 2, nm), Gtau_mid_step(nb, 2, nm), sign_step(nm))
         Gtau_mean_step = 0
         Gtau_mid_step = 0
         sign_step = 0
         do c = 0, n_chains - 1
            if (failed(w2d_ctqmc_get_steps(handle, int(c, c_int), nm, c_loc(ms), c_loc(ds), c_loc(ss)), &
                       "w2d_ctqmc_get_steps")) return
            do s = 1, 2
               do b = 1, nb
                  f = b + nb*(s - 1)
                  Gtau_mean_step(b, s, :) = Gtau_mean_step(b, s, :) + ms(f, :)/beta**2/n_chains
                  Gtau_mid_step(b, s, :) = Gtau_mid_step(b, s, :) + ds(f, :)/beta**2/n_chains
               end do
            end do
            sign_step = sign_step + ss/n_chains
         end do
      end if
      if (w2d_ctqmc_get_gtau_blocks(handle, 0_c_int, nl, c_null_ptr) == 0) then
         allocate (bl(ntau, 2*nb, nl), gtau_blocks(nb, 2, ntau, nl))
         gtau_blocks = 0
         do c = 0, n_chains - 1
            if (failed(w2d_ctqmc_get_gtau_blocks(handle, int(c, c_int), nl, c_loc(bl)), &
                       "w2d_ctqmc_get_gtau_blocks")) return
            do s = 1, 2
               do b = 1, nb
                  f = b + nb*(s - 1)
                  gtau_blocks(b, s, :, :) = gtau_blocks(b, s, :, :) + bl(:, f, :)/n_chains
               end do
            end do
         end do
      end if
   end subroutine collect_steps_and_blocks

   !> MeasGiw: Giw * zScaling / beta (CTQMC.F90:8816) of the sums over the
   !! chains that completed, i.e. sum / (count x mean sign x beta); the
   !! negative frequencies are the complex conjugates
   subroutine collect_giw(cnt, msign)
      real(KINDR), intent(in) :: cnt, msign
      integer(c_int) :: niw
      integer :: b, s, f, n
      real(KINDR), allocatable, target :: g(:,:,:)
      if (allocated(Giw)) deallocate (Giw)
      if (failed(w2d_ctqmc_get_giw(handle, 0_c_int, niw, c_null_ptr), "w2d_ctqmc_get_giw")) return
      if (niw == 0) return
      allocate (g(2, niw, 2*nb), Giw(nb, 2, -niw:niw - 1))
      if (failed(w2d_ctqmc_get_giw(handle, 1_c_int, niw, c_loc(g)), "w2d_ctqmc_get_giw")) return
      do s = 1, 2
         do b = 1, nb
            f = b + nb*(s - 1)
            do n = 0, niw - 1
               Giw(b, s, n) = cmplx(g(1, n + 1, f), g(2, n + 1, f), KINDR)/(cnt*msign*beta)
               Giw(b, s, -n - 1) = conjg(Giw(b, s, n))
            end do
         end do
      end do
   end subroutine collect_giw

   subroutine reset_results()
      if (allocated(Gtau)) deallocate (Gtau, GLeg, occ, DensityMatrix, Histo, OuterSuperstateHisto)
      if (allocated(OuterStateHisto)) deallocate (OuterStateHisto, SignHistoSuperstates, SignHistoStates, &
                                                  TraceContribSuperstates, TraceContribStates, Gtau_full, GLeg_full)
      allocate (Gtau(nb, 2, ntau), GLeg(nb, 2, nleg), occ(nb, 2, nb, 2), DensityMatrix(nstates, nstates))
      allocate (Histo(nb, 2, 0:maxhisto), OuterSuperstateHisto(nsst))
      allocate (OuterStateHisto(nstates), SignHistoSuperstates(nsst), SignHistoStates(nstates))
      allocate (TraceContribSuperstates(nsst), TraceContribStates(nstates))
      allocate (Gtau_full(nb, 2, nb, 2, ntau), GLeg_full(nb, 2, nb, 2, nleg))
      Gtau_full = 0
      GLeg_full = 0
      Gtau = 0
      GLeg = 0
      occ = 0
      DensityMatrix = 0
      Histo = 0
      OuterSuperstateHisto = 0
      mean_sign = 0
      AccAdd = 0
      AccRem = 0
      AccGlob = 0
      AccShift = 0
      TryAdd = 0
      TryRem = 0
      TryGlob = 0
      TryShift = 0
   end subroutine reset_results

   !> get_mc_config_scalars (CTQMC.F90:9358) of chain config_chain
   subroutine get_mc_config_scalars(N, outer_sst, outer_state)
      integer, intent(out) :: N, outer_sst, outer_state
      integer(c_int) :: nn, osst
      real(KINDR) :: tau(1), w(5)
      integer(c_int) :: idum(1)
      N = 0
      outer_sst = 0
      outer_state = 0
      if (failed(w2d_ctqmc_get_config(handle, int(config_chain, c_int), 0_c_int, nn, tau, idum, idum, &
                                      idum, osst, w), "w2d_ctqmc_get_config")) return
      N = nn
      outer_sst = osst
   end subroutine get_mc_config_scalars

   !> get_mc_config (CTQMC.F90:9366): orbitals and spins 1-based, ca 1 = creator
   subroutine get_mc_config(N, taus, orbs, spins, cas, hashybs, outer_sst, outer_state)
      integer, intent(in) :: N
      real(KINDR), dimension(N), intent(out) :: taus
      integer, dimension(N), intent(out) :: orbs, spins, cas, hashybs
      !f2py depend(N) taus, orbs, spins, cas, hashybs
      integer, intent(out) :: outer_sst, outer_state
      integer(c_int) :: nn, osst
      integer(c_int) :: o(max(N, 1)), sp(max(N, 1)), ca(max(N, 1))
      real(KINDR) :: t(max(N, 1)), w(5)
      outer_state = 0
      outer_sst = 0
      if (failed(w2d_ctqmc_get_config(handle, int(config_chain, c_int), int(N, c_int), nn, t, o, sp, &
                                      ca, osst, w), "w2d_ctqmc_get_config")) return
      if (nn /= N) then
         aborted = 2
         return
      end if
      taus = t(1:N)
      orbs = o(1:N) + 1
      spins = sp(1:N) + 1
      cas = ca(1:N)
      hashybs = -1
      outer_sst = osst
   end subroutine get_mc_config

   !> set_mc_config (CTQMC.F90:9378) for chain config_chain
   subroutine set_mc_config(N, taus, orbs, spins, cas, hashybs, outer_sst, outer_state)
      integer, intent(in) :: N
      real(KINDR), dimension(N), intent(in) :: taus
      integer, dimension(N), intent(in) :: orbs, spins, cas, hashybs
      ! N first and required, as cthyb_solver.py passes it (CtHybConfig.set_to_ctqmc)
      !f2py depend(N) taus, orbs, spins, cas, hashybs
      integer, intent(in) :: outer_sst, outer_state
      integer(c_int) :: o(max(N, 1)), sp(max(N, 1)), ca(max(N, 1))
      real(KINDR) :: t(max(N, 1)), w(5)
      if (any(hashybs /= -1)) then
         write (0, '(A)') "MCTQMC: set_mc_config supports hybridised operators only"
         aborted = 2
         return
      end if
      t(1:N) = taus
      o(1:N) = orbs - 1
      sp(1:N) = spins - 1
      ca(1:N) = cas
      if (failed(w2d_ctqmc_set_config(handle, int(config_chain, c_int), int(N, c_int), t, o, sp, ca, &
                                      int(outer_sst, c_int), w), "w2d_ctqmc_set_config")) return
      if (.false.) print *, outer_state
   end subroutine set_mc_config

   !> set_empty_mc_config (CTQMC.F90:9389)
   subroutine set_empty_mc_config(outer_sst, outer_state)
      integer, intent(in) :: outer_sst, outer_state
      real(KINDR) :: t(0)
      integer :: i(0)
      call set_mc_config(0, t, i, i, i, i, outer_sst, outer_state)
   end subroutine set_empty_mc_config

   !> dest_ctqmc (CTQMC.F90:901)
   subroutine dest_ctqmc()
      integer(c_int) :: st
      if (c_associated(handle)) then
         st = w2d_ctqmc_destroy(handle)
         handle = c_null_ptr
      end if
   end subroutine dest_ctqmc

end module MCTQMC
