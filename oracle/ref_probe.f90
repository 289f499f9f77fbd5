! Probe driver for the reference THERMACOND routines (test infrastructure, this container
! only).  It runs the job=1 sequence of THERMACOND/kappa.f90 on a single process -- the
! preamble of kappa.f90:178-351 without the band-structure, Grueneisen and DOS
! diagnostics, then kappa.f90:389-871 with ksubset = [1, nibz] and no MPI calls -- and
! dumps every intermediate array as raw little-endian binary into ./dump/ so the oracle
! restatement and the GPU path can be checked against the reference itself.
!
! Extra input: probe_temps.txt (optional) = list of temperatures in K, one per line; the
! temperature-dependent part (kappa.f90:557-856) is repeated for each.  Without it the
! single temperature tmin is used, as kappa.f90:565 does.
!
! probe_selfq.txt (optional): nq, then nq points q (Cartesian, the units of kpc), one per
! line: the three-phonon self-energy of jobs 2/6/7 (kappa.f90:900-925) at every branch of
! each q, function_self_w3 (rest.f90:4012-4124) at omega = omega(q,la) from get_freq and
! T = tmin, dumped as selfE_omega (ndyn,nq), selfE_n and selfE_u (complex (ndyn,nq)), and the
! Gaussian form function_self_w4 (rest.f90:4126-4240) as selfE4_n and selfE4_u; with
! REF_PROBE_SELF_ONLY set the probe stops there.
!
! REF_PROBE_ITER=n (environment, optional): after kappa_RTA, also run the iterative BTE of
! job 3 (kappa.f90:1373-1487: update_mpi, F <- F + update_mix*(F2 - F), mode_kappa, stop when
! the kappa change norm is below 1e-4) for at most min(n, max_iter) iterations, dumping the
! first update F2, kappa and tauinv_eff(nibz,ndyn) after every iteration and the final F
! (mode_kappa clears tauinv_eff for every mode, so no other element survives); F is then reset to F_RTA so
! that the job-1 CG below starts from the same guess as before.
!
! probe_calg.txt (optional): n, then n element names, one per line: cal_g (sub.f90:3766-4693)
! for each, dumped as calg (n); the probe stops there.
!
! REF_PROBE_ISO (environment, optional): isotope scattering, iso = 1, routine by routine:
! cal_giso and P_iso_split (kappa.f90:522-543) after the FGR, dumped as giso and Piso; the
! P1sm/P2sm blocks written as v33sq_delta.NNN.dat (the files calculate_tauRTA_iso reads);
! per temperature calculate_tauRTA_iso after calculate_tauRTA_mpi2 (kappa.f90:574-578),
! update_iso in place of update_mpi (kappa.f90:1401-1407) and CollisionMatrix_iso in place of
! CollisionMatrix_mpi (kappa.f90:710-716), its Collision_Matrix then feeding RHS, CG and
! mode_kappa_d as before.  (kappa.f90's own job-1 flow gathers Qvaluempi after the iso
! branch, which calculate_tauRTA_iso never sets; the probe keeps the routine's Qvalue.)
program ref_probe
 use ios
 use kpoints
 use om_dos
 use atoms_force_constants
 use phi3
 use phi3_sy
 use eigen
 use params
 use constants
 use born
 use exactBTE2
 use mod_ksubset2
 use tetrahedron
 use lattice
 use svd_stuff
 implicit none
 integer :: i, j, k, s, ksubset(2), nv3_split, nv3o, ksub, it, ntemps, ios_
 integer :: is, js, ks, n1, l1, n, narms, kvecop(48)
 real(8), allocatable :: dkc(:), temps(:), kpc_mesh(:,:)
 real(8) :: tempk, fm, kvecstar(3,48)
 integer, allocatable :: tetc(:,:), starn(:), starop(:,:), starn0(:,:)
 character(len=64) :: tag
 character(len=256) :: envs
 integer(8) :: c0, c1, crate
 logical :: ex
 integer :: maxit, iter_, niter, nsq, la
 real(8), allocatable :: selfq(:,:), seom(:,:), eivq(:)
 complex(8), allocatable :: sen(:,:), seu(:,:), evq(:,:), sen4(:,:), seu4(:,:)
 complex(8) :: nself, uself
 real(8) :: tcm
 real(8) :: normk
 real(8), allocatable :: kiter(:,:,:,:), teiter(:)
 logical :: iso_
 integer :: i0, i1, n2, l2, l3
 character(len=64) :: fname
 character(len=2) :: elname
 real(8), allocatable :: calg(:)

 const33 = 1d30*(hbar/uma)**1.5d0/(sqrt(ee*1d20/uma))**1.5d0*ee/h_plank/100/c_light
 inquire(file='probe_calg.txt', exist=ex)
 if (ex) then
    open(78, file='probe_calg.txt', status='old')
    read(78,*) n
    allocate(calg(n))
    do i = 1, n
       read(78,'(a)') elname
       call cal_g(elname, calg(i))
    enddo
    close(78)
    call system('mkdir -p dump')
    call dumpr('calg', calg, n)
    stop
 endif
 call system('mkdir -p dump')

 call read_params
 call read_input_fit
 call read_lattice
 call read_born
 nkc = nc(1)*nc(2)*nc(3)
 call read_fc234
 ndyn = 3*natoms0

 allocate(kpc(3,nkc), mappos(nkc), wk(nkc))
 call make_kp_reg(nc,g1,g2,g3,shft,kpc,wk)
 allocate(kpc_mesh(3,nkc)); kpc_mesh = kpc
 call get_negatives(nkc,kpc,mappos,npos)
 call get_weights(nkc,kpc)
 call allocate_eig(ndyn,nkc,nkc)
 allocate(dkc(nkc)); dkc = 0
 call get_frequencies(nkc,kpc,dkc,ndyn,eigenval,ndyn,eigenvec,ucors,veloc)
 call dumpc('eigenvec_raw', eigenvec, size(eigenvec))
 call subst_eivecs(ndyn,nkc,eigenvec,kpc,npos,mappos)
 call allocate_tetra(nc(1),nc(2),nc(3),wmesh,ndyn)
 call make_kp_reg_tet(nc(1),nc(2),nc(3),shft(1),shft(2),shft(3),kpc,wk)

 call system('mkdir -p dump')
 open(77, file='dump/dims.txt', status='replace')
 write(77,*) nkc, nibz, ndyn, natoms0, nterms(2), nterms(3), lattpgcount, npos
 close(77)
 call dumpr('kpc_mesh', kpc_mesh, 3*nkc)
 call dumpr('kpc', kpc, 3*nkc)
 call dumpi('mapibz', mapibz, nkc)
 call dumpi('mapinv', mapinv, nibz)
 call dumpi('mappos', mappos, nkc)
 call dumpr('kibz', kibz, 3*nibz)
 call dumpr('wibz', wibz, nibz)
 call dumpr('eigenval', eigenval, ndyn*nkc)
 call dumpc('eigenvec', eigenvec, ndyn*ndyn*nkc)
 call dumpr('veloc', veloc, 3*ndyn*nkc)
 call dumpr('op_kmatrix', op_kmatrix, 9*lattpgcount)
 allocate(tetc(4,6*nkc))
 do i = 1, 6*nkc
    do j = 1, 4
       tetc(j,i) = tet(i)%p(j)%i
    enddo
 enddo
 call dumpi('tet', tetc, 24*nkc)
 ! star lists on the mesh (kstarsinfbz1) for every IBZ point
 allocate(starn(nibz), starop(48,nibz), starn0(48,nibz))
 starop = 0; starn0 = 0
 do i = 1, nibz
    call kstarsinfbz1(kibz(:,i),nkc,kpc,narms,kvecstar,kvecop)
    starn(i) = narms
    do s = 1, narms
       starop(s,i) = kvecop(s)
       call get_k_info(kvecstar(:,s),nc,n,is,js,ks,g1,g2,g3,k)
       starn0(s,i) = n
    enddo
 enddo
 call dumpi('star_n', starn, nibz)
 ! ---------------- self-energy at given q (jobs 2/6/7) ----------------
 inquire(file='probe_selfq.txt', exist=ex)
 if (ex) then
    open(78, file='probe_selfq.txt', status='old')
    read(78,*) nsq
    allocate(selfq(3,nsq), seom(ndyn,nsq), sen(ndyn,nsq), seu(ndyn,nsq), eivq(ndyn), evq(ndyn,ndyn))
    allocate(sen4(ndyn,nsq), seu4(ndyn,nsq))
    do i = 1, nsq
       read(78,*) selfq(:,i)
    enddo
    close(78)
    tcm = tmin*k_b/(100*h_plank*c_light)
    do i = 1, nsq
       call get_freq(selfq(:,i), ndyn, eivq, evq)
       do la = 1, ndyn
          seom(la,i) = sqrt(abs(eivq(la)))*cnst
          call function_self_w3(selfq(:,i), la, seom(la,i), tcm, nself, uself)
          sen(la,i) = nself
          seu(la,i) = uself
          call function_self_w4(selfq(:,i), la, seom(la,i), tcm, nself, uself)
          sen4(la,i) = nself
          seu4(la,i) = uself
       enddo
    enddo
    call dumpr('selfE_omega', seom, ndyn*nsq)
    call dumpc('selfE_n', sen, ndyn*nsq)
    call dumpc('selfE_u', seu, ndyn*nsq)
    call dumpc('selfE4_n', sen4, ndyn*nsq)
    call dumpc('selfE4_u', seu4, ndyn*nsq)
    call get_environment_variable('REF_PROBE_SELF_ONLY', envs, status=ios_)
    if (ios_ == 0) stop
 endif
 call dumpi('star_op', starop, 48*nibz)
 call dumpi('star_n0', starn0, 48*nibz)

 ! ---------------- job 1: V3 and FGR on the whole IBZ (one "rank") ----------------
 ksubset(1) = 1; ksubset(2) = nibz
 ! timing mode: REF_PROBE_V3_SAMPLE="lo hi" runs only calculate_w3_ibz_split_sq on that
 ! IBZ block and prints its wall time and element count
 call get_environment_variable('REF_PROBE_V3_SAMPLE', envs, status=ios_)
 if (ios_ == 0 .and. len_trim(envs) > 0) then
    read(envs,*) ksubset(1), ksubset(2)
    ksub = ksubset(2) - ksubset(1) + 1
    nv3_split = nkc*ksub*ndyn**3
    call allocate_v33sq(nv3_split)
    allocate(v33s8(ksub,nkc,ndyn,ndyn,ndyn))
    call system_clock(c0, crate)
    call calculate_w3_ibz_split_sq(ksubset,nv3_split,ndyn,nkc,kpc,eigenval,eigenvec,nv3o)
    call system_clock(c1)
    write(*,'(a,i0,a,es16.8)') 'REF_PROBE_V3 elements=', nv3_split, ' seconds=', dble(c1-c0)/dble(crate)
    stop
 endif
 ! block mode: REF_PROBE_SUBSET="lo hi" runs V3, the FGR and calculate_tauRTA_mpi2 on that IBZ
 ! block only, as one rank of kappa.f90 does with its ksubset (kappa.f90:420-431), dumps
 ! v33s8, P1sm, P2sm and per temperature the block's Qvaluempi/_N/_U (ksub x ndyn) as
 ! T<t>_Qsub, T<t>_QNsub, T<t>_QUsub, and stops
 call get_environment_variable('REF_PROBE_SUBSET', envs, status=ios_)
 if (ios_ == 0 .and. len_trim(envs) > 0) then
    read(envs,*) ksubset(1), ksubset(2)
    ksub = ksubset(2) - ksubset(1) + 1
    nv3_split = nkc*ksub*ndyn**3
    call allocate_v33sq(nv3_split)
    allocate(v33s8(ksub,nkc,ndyn,ndyn,ndyn))
    call calculate_w3_ibz_split_sq(ksubset,nv3_split,ndyn,nkc,kpc,eigenval,eigenvec,nv3o)
    call dumpr('v33s8', v33s8, size(v33s8))
    allocate(P1smpi(ksub,nkc,ndyn,ndyn,ndyn), P2smpi(ksub,nkc,ndyn,ndyn,ndyn))
    P1smpi = 0; P2smpi = 0
    call calculate_FGR_mpi2(ksubset,ndyn,nv3o)
    call dumpr('P1sm', P1smpi, size(P1smpi))
    call dumpr('P2sm', P2smpi, size(P2smpi))
    call dumpr('frequency', frequency, nkc*ndyn)
    allocate(Qvaluempi(ksub,ndyn), Qvalue_Nmpi(ksub,ndyn), Qvalue_Umpi(ksub,ndyn))
    open(78, file='probe_temps.txt', status='old')
    do
       read(78,*,iostat=ios_) tempk
       if (ios_ /= 0) exit
       write(tag,'(a,i0)') 'T', nint(tempk)
       call allocate_FGR(ksub,nkc,ndyn,ndyn,ndyn)
       call distribution(nkc,ndyn,tempk)
       call calculate_tauRTA_mpi2(ksubset,ndyn,tempk)
       call dumpr(trim(tag)//'_Qsub', Qvaluempi, ksub*ndyn)
       call dumpr(trim(tag)//'_QNsub', Qvalue_Nmpi, ksub*ndyn)
       call dumpr(trim(tag)//'_QUsub', Qvalue_Umpi, ksub*ndyn)
    enddo
    close(78)
    write(*,'(a,i0,a,i0)') 'REF_PROBE_SUBSET done ', ksubset(1), ' ', ksubset(2)
    stop
 endif
 ksub = nibz
 nv3_split = nkc*ksub*ndyn**3
 call allocate_v33sq(nv3_split)
 allocate(v33s8(ksub,nkc,ndyn,ndyn,ndyn))
 call calculate_w3_ibz_split_sq(ksubset,nv3_split,ndyn,nkc,kpc,eigenval,eigenvec,nv3o)
 call dumpr('v33s8', v33s8, size(v33s8))
 allocate(P1smpi(ksub,nkc,ndyn,ndyn,ndyn), P2smpi(ksub,nkc,ndyn,ndyn,ndyn))
 P1smpi = 0; P2smpi = 0
 call calculate_FGR_mpi2(ksubset,ndyn,nv3o)
 call dumpr('P1sm', P1smpi, size(P1smpi))
 call dumpr('P2sm', P2smpi, size(P2smpi))
 call dumpr('frequency', frequency, nkc*ndyn)

 call get_environment_variable('REF_PROBE_ISO', envs, status=ios_)
 iso_ = (ios_ == 0)
 if (iso_) then
    allocate(giso(natoms0), Pisos(nibz,ndyn))
    call cal_giso(giso)
    call P_iso_split(ksubset,giso,ndyn,eigenvec)
    call dumpr('giso', giso, natoms0)
    call dumpr('Piso', Pisos, nibz*ndyn)
    do n = 1, (nibz + ksub_size - 1)/ksub_size
       i0 = (n - 1)*ksub_size + 1
       i1 = min(n*ksub_size, nibz)
       write(fname,'(a,i3.3,a)') 'v33sq_delta.', n, '.dat'
       open(80, file=trim(v3path)//trim(fname), form='unformatted', status='replace')
       write(80) (i1 - i0 + 1)*nkc*ndyn**3
       do n1 = i0, i1
          do l1 = 1, ndyn
             do n2 = 1, nkc
                do l2 = 1, ndyn
                   do l3 = 1, ndyn
                      write(80) P1smpi(n1,n2,l1,l2,l3), P2smpi(n1,n2,l1,l2,l3)
                   enddo
                enddo
             enddo
          enddo
       enddo
       close(80)
    enddo
 endif

 ntemps = 1
 allocate(temps(64)); temps(1) = tmin
 inquire(file='probe_temps.txt', exist=ex)
 if (ex) then
    open(78, file='probe_temps.txt', status='old')
    ntemps = 0
    do
       read(78,*,iostat=ios_) tempk
       if (ios_ /= 0) exit
       ntemps = ntemps + 1
       temps(ntemps) = tempk
    enddo
    close(78)
 endif

 allocate(FFF(nibz,ndyn,3), Qvaluempi(ksub,ndyn), Qvalue_Nmpi(ksub,ndyn), Qvalue_Umpi(ksub,ndyn))
 allocate(Collision_Matrixmpi(ksub*ndyn*3,nibz*ndyn*3), RHS(nibz*ndyn*3), DirectSolution(nibz*ndyn*3))
 do it = 1, ntemps
    tempk = temps(it)
    write(tag,'(a,i0)') 'T', nint(tempk)
    call allocate_FGR(ksub,nkc,ndyn,ndyn,ndyn)
    call distribution(nkc,ndyn,tempk)
    call calculate_tauRTA_mpi2(ksubset,ndyn,tempk)
    Qvalue = Qvaluempi; Qvalue_N = Qvalue_Nmpi; Qvalue_U = Qvalue_Umpi
    if (iso_) then
       Piso = Pisos
       call calculate_tauRTA_iso(ndyn,tempk)
    endif
    call Qandtau(ndyn,tempk)
    call calculate_RTA_distributionfunction(nibz,nkc,kibz,ndyn,tempk,veloc,eigenval)
    call mode_kappa(ndyn,tempk,veloc,eigenval)
    open(3450,file='tau_RTA_'//trim(tag)//'.dat',status='unknown')
    open(6450,file='kappa_RTA_'//trim(tag)//'.dat',status='unknown')
    call write_kappa(ndyn,tempk,veloc,eigenval,6450,3450)
    close(3450); close(6450)
    call dumpr(trim(tag)//'_dist', dist, nkc*ndyn)
    call dumpr(trim(tag)//'_Q', Qvalue, nibz*ndyn)
    call dumpr(trim(tag)//'_QN', Qvalue_N, nibz*ndyn)
    call dumpr(trim(tag)//'_QU', Qvalue_U, nibz*ndyn)
    call dumpr(trim(tag)//'_tauN', tauinv_N, nibz*ndyn)
    call dumpr(trim(tag)//'_tauU', tauinv_U, nibz*ndyn)
    call dumpr(trim(tag)//'_tautot', tauinv_tot, nibz*ndyn)
    call dumpr(trim(tag)//'_taueff', tauinv_eff, nibz*ndyn)
    call dumpr(trim(tag)//'_FRTA', FFF, nibz*ndyn*3)
    call dumpr(trim(tag)//'_kappaRTA', kappa, ndyn*9)
    maxit = 0
    call get_environment_variable('REF_PROBE_ITER', envs, status=ios_)
    if (ios_ == 0) read(envs,*) maxit
    maxit = min(maxit, max_iter)
    if (maxit > 0) then
       allocate(F_RTA(nibz,ndyn,3), F2(nibz,ndyn,3), F2mpi(ksub,ndyn,3))
       allocate(kappa_old(ndyn,3,3), kappa_new(ndyn,3,3), e(ndyn,3,3), kiter(ndyn,3,3,maxit), teiter(maxit))
       F_RTA = FFF
       niter = 0
       do iter_ = 1, maxit
          kappa_old = kappa
          if (iso_) then
             call update_iso(ndyn)
          else
             call update_mpi(ndyn,ksubset)
             F2 = F2mpi
          endif
          if (iter_ == 1) call dumpr(trim(tag)//'_F2first', F2, nibz*ndyn*3)
          FFF(:,:,:) = FFF(:,:,:) + update_mix*(F2(:,:,:)-FFF(:,:,:))
          call mode_kappa(ndyn,tempk,veloc,eigenval)
          kappa_new = kappa
          niter = iter_
          kiter(:,:,:,iter_) = kappa
          teiter(iter_) = tauinv_eff(nibz,ndyn)
          e = (kappa_new-kappa_old)**2/(ndyn*3*3)
          normk = sqrt(sum(e(:,:,:)))
          if (normk .lt. 0.0001) exit
       enddo
       call dumpr(trim(tag)//'_kappaIter', kiter, ndyn*9*niter)
       call dumpr(trim(tag)//'_Fiter', FFF, nibz*ndyn*3)
       call dumpr(trim(tag)//'_taueffIter', teiter, niter)
       FFF = F_RTA
       deallocate(F_RTA, F2, F2mpi, kappa_old, kappa_new, e, kiter, teiter)
    endif
    allocate(Collision_Matrix(nibz*ndyn*3,nibz*ndyn*3))
    if (iso_) then
       call CollisionMatrix_iso(ndyn)
       call dumpr(trim(tag)//'_CM', Collision_Matrix, size(Collision_Matrix))
    else
       call CollisionMatrix_mpi(ksubset,ndyn)
       call dumpr(trim(tag)//'_CM', Collision_Matrixmpi, size(Collision_Matrixmpi))
       Collision_Matrix = Collision_Matrixmpi
    endif
    call cal_RHS_IBZ(ndyn,nkc,tempk,veloc,eigenval)
    call dumpr(trim(tag)//'_RHS', RHS, nibz*ndyn*3)
    s = 0
    do is = 1, nibz
       do js = 1, ndyn
          do ks = 1, 3
             s = s + 1
             DirectSolution(s) = FFF(is,js,ks)
          enddo
       enddo
    enddo
    call cg_quadratic(nibz*ndyn*3,Collision_Matrix,RHS,DirectSolution,fm,100,2)
    call dumpr(trim(tag)//'_x', DirectSolution, nibz*ndyn*3)
    call mode_kappa_d(ndyn,tempk,veloc,eigenval,DirectSolution)
    call dumpr(trim(tag)//'_kappaBTE', kappa, ndyn*9)
    call dumpr(trim(tag)//'_taueffd', tauinv_eff, nibz*ndyn)
    open(7450,file='tau_RTAandeff-cg_'//trim(tag)//'.dat',status='unknown')
    open(8450,file='kappa_direct-cg_'//trim(tag)//'.dat',status='unknown')
    call write_kappa(ndyn,tempk,veloc,eigenval,8450,7450)
    close(7450); close(8450)
    deallocate(Collision_Matrix)
    call deallocate_FGR
    deallocate(Piso)
 enddo

contains

 subroutine dumpr(name, a, n)
  character(len=*), intent(in) :: name
  integer, intent(in) :: n
  real(8), intent(in) :: a(n)
  open(79, file='dump/'//name//'.bin', access='stream', form='unformatted', status='replace')
  write(79) a
  close(79)
 end subroutine dumpr

 subroutine dumpi(name, a, n)
  character(len=*), intent(in) :: name
  integer, intent(in) :: n
  integer, intent(in) :: a(n)
  open(79, file='dump/'//name//'.bin', access='stream', form='unformatted', status='replace')
  write(79) a
  close(79)
 end subroutine dumpi

 subroutine dumpc(name, a, n)
  character(len=*), intent(in) :: name
  integer, intent(in) :: n
  complex(8), intent(in) :: a(n)
  open(79, file='dump/'//name//'.bin', access='stream', form='unformatted', status='replace')
  write(79) a
  close(79)
 end subroutine dumpc

end program ref_probe
