! Test-infrastructure driver for the reference Fortran (never part of the product).
!
! Builds the Yin-Yang tree and blocks with the reference's own modules (compiled
! in place from 3_6/ by oracle/Makefile), loads a background and initial state
! written by tests/golden/make_golden.py, and runs the reference's own driver
! sequence (3_6/test1.f90:110-148 and :56-58):
!   ModCommunication_SetGCAll, state, two ModCommunication_SendRecvGC_new(.false.),
!   then ModAdvance_rk4 per step,
! with the reference's ModCommunication.f90 and ModAdvance.f90 themselves, over
! MPI.  Run under `mpirun -n N` the blocks split over ranks exactly as
! ModAllocation_Init assigns them; the output file is the same for every N.
! The background replication follows ModBlock_InitSSM (3_6/ModBlock.f90:251-264);
! the 1-D profiles come from the input file because the FreeEOS table that
! ModStratification_new reads is not in the reference tree.
!
! Built with -DFALCON_DROPIN (oracle/Makefile target `dropin`), the same driver
! links the Fortran shims of falcon_amd/fortran in place of the reference's
! ModCommunication / ModAdvance / ModCheck, stepping on the GPU (one rank).
!
! Usage: [mpirun -n N] falcon_ref_harness <input.bin> <output.bin>
! The input's state count is 8, or 9 for a run with GLM divergence cleaning
! (DivB_option=1, 3_6/ModAdvance.f90:82-84): psi is the ninth variable of the
! initial state and of the state dumps.
! Input modes: 1 tables only; 2 tables, dumps and steps; 3 steps only (timing:
! no output beyond the file header and the stepping loop's wall time); 4 the
! reference's ModSave_AllCells output of the exchanged initial state, written
! into the working directory (AllCells_00000000/), and no steps.
program falcon_ref_harness
    use MPI
    use ModParameters,  only:   ni,nj,nk,ng,nvar,iGeometry,iEquation,MpiSize,MpiRank,&
                                if_involve_B,Artificial_heating_ratio,CFL,rLevelInitial,&
                                nAMRs,AMRs,DivB_option
    use ModControl,     only:   nMaxBlocksPerRank
    use ModBlock,       only:   BlockType,ModBlock_InitGrid,ModBlock_InitPrimitives
    use ModYinYangTree, only:   YYTree,YinYangTree_InitTree,YinYangTree_LabelNodes,&
                                YinYangTree_GetLocalBlocksInfo,YinYangTree_if_top_bottom
    use ModAllocation,  only:   ModAllocation_Init,nLocalLeafNodes,iLocalLeafNodes
    use ModAMR,         only:   ModAMR_set_grid
    use ModEquation,    only:   ModEquation_Dynamo
    use ModDiffusion,   only:   ModDiffusion_Aritificial_1
    use ModBoundary,    only:   ModBoundary_Dynamo_primitives
    use ModCommunication, only: ModCommunication_SetGCAll,ModCommunication_SendRecvGC_new
    use ModAdvance,       only: ModAdvance_rk4
    use ModCheck,         only: ModCheck_primitive
    use ModSavePlot,      only: ModSave_AllCells
#ifdef FALCON_DROPIN
    use ModFalconShim,    only: falcon_sync_to_host
#endif
    implicit none

    type(YYTree),target         ::  Tree
    character(len=512)          ::  fin,fout
    integer                     ::  ui,iamr,iflag(3),iglob,ib,jb,nsteps,ndump,mode,istep,idump,ierr
    integer,allocatable         ::  dump_steps(:)
    real(8)                     ::  r_range_cgs(2),rng(3,2),dt,t0,t1
    real(8),allocatable         ::  rtp_ranges(:,:,:),bg(:,:),st(:,:,:,:)
    integer                     ::  nstate,v
    type(BlockType),pointer     ::  B

    call MPI_INIT(ierr)
    call MPI_Comm_size(MPI_COMM_WORLD,MpiSize,ierr)
    call MPI_Comm_rank(MPI_COMM_WORLD,MpiRank,ierr)

    call get_command_argument(1,fin)
    call get_command_argument(2,fout)
    open(newunit=ui,file=trim(fin),access='stream',form='unformatted',status='old',action='read')

    read(ui) ni,nj,nk,ng,nstate
    read(ui) r_range_cgs
    read(ui) nAMRs
    allocate(AMRs(max(nAMRs,1)))
    do iamr=1,nAMRs
        read(ui) iflag,iglob,rng
        AMRs(iamr)%iLevel=iamr
        AMRs(iamr)%rtp_if_divide=(iflag/=0)
        AMRs(iamr)%if_global=(iglob/=0)
        AMRs(iamr)%rtp_range_Rsun=rng
    end do
    read(ui) CFL
    read(ui) nsteps
    read(ui) ndump
    allocate(dump_steps(max(ndump,1)))
    if (ndump>0) read(ui) dump_steps(1:ndump)
    read(ui) mode

    ! Global parameters as the reference's #GRID/#EQUATION setup leaves them.
    iGeometry=1; iEquation=1; if_involve_B=.true.; DivB_option=merge(1,0,nstate==9)
    Artificial_heating_ratio=1.0d0
    nvar=9
    rLevelInitial=0

    ! Tree (3_6/test1.f90:113-119 with YinYangTree_SetAll's steps split out so the
    ! background can come from the input file instead of ModStratification_new).
    call YinYangTree_InitTree(Tree,r_range_cgs)
    call ModAMR_set_grid(Tree)
    Tree%NumLeafNodes=0
    Tree%NumLeafNodes_YinYang=0
    call YinYangTree_LabelNodes(Tree,Tree%Yin)
    Tree%NumLeafNodes_YinYang(1)=Tree%NumLeafNodes
    call YinYangTree_LabelNodes(Tree,Tree%Yang)
    Tree%NumLeafNodes_YinYang(2)=Tree%NumLeafNodes-Tree%NumLeafNodes_YinYang(1)
    call ModAllocation_Init(Tree%NumLeafNodes)
    ! the message tags of ModCommunication scale with nMaxBlocksPerRank
    ! (3_6/ModCommunication.f90:947); raised, as a #MPI parameter would, to
    ! hold every rank's share (3_6/test1.f90:122-130)
    nMaxBlocksPerRank=max(nMaxBlocksPerRank,(Tree%NumLeafNodes+MpiSize-1)/MpiSize)
    allocate(rtp_ranges(3,2,nLocalLeafNodes))
    allocate(Tree%LocalBlocks(nLocalLeafNodes))
    allocate(Tree%iBlocks(nLocalLeafNodes))
    Tree%iBlocks=iLocalLeafNodes
    Tree%nLocalBlocks=nLocalLeafNodes
    ib=1
    call YinYangTree_GetLocalBlocksInfo(Tree,Tree%Yin ,rtp_ranges,ib)
    call YinYangTree_GetLocalBlocksInfo(Tree,Tree%Yang,rtp_ranges,ib)

    allocate(bg(-ng+1:ni+ng,8),st(-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng,nstate))

    do ib=1,Tree%nLocalBlocks
        B=>Tree%LocalBlocks(ib)
        ! ModBlock_Init (3_6/ModBlock.f90:115-186), itype_equation=1, without InitSSM
        B%iBlock=Tree%iBlocks(ib)
        B%if_yin=(Tree%iBlocks(ib).le.Tree%NumLeafNodes_YinYang(1))
        B%itype_equation=1
        B%if_involve_B=.true.
        B%if_bottom=.false.; B%if_top=.false.
        call ModBlock_InitGrid(B,rtp_ranges(:,:,ib))
        B%nvar=9
        B%rho1_=1; B%vr_=2; B%vt_=3; B%vp_=4; B%br_=5; B%bt_=6; B%bp_=7; B%s1_=8; B%psi_=9
        call ModBlock_InitPrimitives(B,.true.)
        B%if_initialized=.true.
        call YinYangTree_if_top_bottom(B,Tree%r_range_CGS)
    end do

    ! 1-D background of every local block: read before ModCommunication_SetGCAll,
    ! as YinYangTree_SetAll's ModBlock_InitSSM sets it before test1.f90:133
    if (mode>=2) then
        do jb=1,Tree%NumLeafNodes
            read(ui) bg
            read(ui) st
            ib=jb-Tree%iBlocks(1)+1
            if (ib>=1 .and. ib<=Tree%nLocalBlocks) call harness_background(Tree%LocalBlocks(ib),bg)
        end do
    end if

    ! Tables (3_6/test1.f90:133)
    call ModCommunication_SetGCAll(Tree)

    ! ---- output: layout and tables
    if (MpiRank==0) then
        block
            integer :: uo
            open(newunit=uo,file=trim(fout),access='stream',form='unformatted',status='replace')
            write(uo) Tree%NumLeafNodes,Tree%NumLeafNodes_YinYang(1)
            close(uo)
        end block
    end if
    call emit(1,0,0.0d0)
    call emit(2,0,0.0d0)
    call emit(3,0,0.0d0)
    call emit(4,0,0.0d0)

    if (mode>=2) then
        ! initial state of every local block: ModInitiation_DoAll's place (test1.f90:140)
        rewind(ui)
        call skip_header()
        do jb=1,Tree%NumLeafNodes
            read(ui) bg
            read(ui) st
            ib=jb-Tree%iBlocks(1)+1
            if (ib>=1 .and. ib<=Tree%nLocalBlocks) then
                B=>Tree%LocalBlocks(ib)
                B%primitive_IV=0.0d0
                do v=1,min(nstate,9)
                    B%primitive_IV(:,:,:,v)=st(:,:,:,v)
                end do
                B%primitive_rk_IV=B%primitive_IV
                B%primitive=>B%primitive_IV
            end if
        end do

        ! initial ghost fill (3_6/test1.f90:143-144)
        call ModCommunication_SendRecvGC_new(Tree,.false.)
        call ModCommunication_SendRecvGC_new(Tree,.false.)
#ifdef FALCON_DROPIN
        call falcon_sync_to_host(Tree)
#endif
        call emit(5,0,0.0d0)
        if (mode==4) then
            call ModSave_AllCells(Tree,0,77)
            nsteps=0
        end if

        ! single-operator dumps on the exchanged initial state
        if (mode==2) then
        do ib=1,Tree%nLocalBlocks
            B=>Tree%LocalBlocks(ib)
            B%primitive=>B%primitive_IV
            call ModBoundary_Dynamo_primitives(B)
            B%EQN_update_R_IV=0.0d0
            call ModEquation_Dynamo(B)
        end do
        call emit(6,0,0.0d0)
        do ib=1,Tree%nLocalBlocks
            B=>Tree%LocalBlocks(ib)
            B%EQN_update_R_IV=0.0d0
            call ModDiffusion_Aritificial_1(B,B%EQN_update_R_IV,2,.false.)
        end do
        call emit(6,0,0.0d0)
        ! restore the state the stepping starts from (BCs above touched ghosts,
        ! which ModAdvance_rk4 overwrites at stage 1 anyway)
        do ib=1,Tree%nLocalBlocks
            Tree%LocalBlocks(ib)%EQN_update_R_IV=0.0d0
        end do
        end if

        idump=1
        call MPI_BARRIER(MPI_COMM_WORLD,ierr)
        t0=MPI_Wtime()
        do istep=1,nsteps
            ! 3_6/test1.f90:58,61 (DoCheck)
            call ModAdvance_rk4(Tree,.true.,dt)
            call ModCheck_primitive(Tree,.false.)
            if (idump<=ndump) then
                if (dump_steps(idump)==istep) then
                    ! the dumps stay outside the timed loop
                    t1=MPI_Wtime()
#ifdef FALCON_DROPIN
                    call falcon_sync_to_host(Tree)
#endif
                    call emit(5,istep,dt)
                    idump=idump+1
                    t0=t0+(MPI_Wtime()-t1)
                end if
            end if
        end do
        call MPI_BARRIER(MPI_COMM_WORLD,ierr)
        t1=MPI_Wtime()
        ! wall time of the stepping loop (the CPU baseline reads this line)
        if (MpiRank==0) write(*,'(a,i0,a,es16.9,a,i0)') 'harness_steps ',nsteps,' seconds ',t1-t0,' ranks ',MpiSize
    end if
    close(ui)
    call MPI_FINALIZE(ierr)

contains

    subroutine skip_header()
        integer :: h5(5),k,n1,n2,n3,md,ia(4)
        real(8) :: r2(2),r6(6),c
        read(ui) h5
        read(ui) r2
        read(ui) n1
        do k=1,n1
            read(ui) ia,r6
        end do
        read(ui) c
        read(ui) n2
        read(ui) n3
        if (n3>0) read(ui) dump_steps(1:n3)
        read(ui) md
    end subroutine skip_header

    ! One section of the output, written rank after rank (each rank's blocks are a
    ! contiguous, ordered range of leaf labels, ModAllocation.f90:25-61), so the
    ! file lists every block in label order whatever the rank count.
    ! what: 1 layout, 2 GC_iBlocks_III, 3 GC_targets, 4 GC_sources, 5 state dump
    ! (step, dt, primitive_IV), 6 EQN_update_R_IV.
    subroutine emit(what,istep,dt)
        integer,intent(in)          ::  what,istep
        real(8),intent(in)          ::  dt
        integer                     ::  r,uo,ib,it
        type(BlockType),pointer     ::  B
        if (mode==3) return
        do r=0,MpiSize-1
            if (r==MpiRank) then
                open(newunit=uo,file=trim(fout),access='stream',form='unformatted',status='old',position='append')
                if (what==5 .and. MpiRank==0) write(uo) istep,dt
                do ib=1,Tree%nLocalBlocks
                    B=>Tree%LocalBlocks(ib)
                    select case(what)
                    case(1)
                        write(uo) B%xijk_range
                        write(uo) merge(1,0,B%if_yin),merge(1,0,B%if_top),merge(1,0,B%if_bottom)
                        write(uo) B%xi_I,B%xj_I,B%xk_I
                    case(2)
                        write(uo) B%GC_iBlocks_III
                    case(3)
                        write(uo) B%nGC_targets
                        do it=1,B%nGC_targets
                            write(uo) B%GC_targets(it)%iBlock,B%GC_targets(it)%nGC
                            write(uo) B%GC_targets(it)%ijk_list
                            write(uo) B%GC_targets(it)%xijk_list
                        end do
                    case(4)
                        write(uo) B%nGC_sources
                        do it=1,B%nGC_sources
                            write(uo) B%GC_sources(it)%iBlock,B%GC_sources(it)%nGC
                            write(uo) B%GC_sources(it)%ijk_list
                            write(uo) B%GC_sources(it)%xijk_list
                        end do
                    case(5)
                        write(uo) B%primitive_IV(:,:,:,1:min(nstate,9))
                    case(6)
                        write(uo) B%EQN_update_R_IV(:,:,:,1:8)
                    end select
                end do
                close(uo)
            end if
            call MPI_BARRIER(MPI_COMM_WORLD,ierr)
        end do
    end subroutine emit

    ! Background replication exactly as ModBlock_InitSSM (3_6/ModBlock.f90:252-264).
    ! bg columns: g, rho0, p0, T0, gamma1, gamma3, heat, xi_rsst.
    subroutine harness_background(B,bg)
        type(BlockType),target      ::  B
        real(8),intent(in)          ::  bg(-ng+1:ni+ng,8)
        integer                     ::  j,k
        allocate(B%gamma1_III        (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%gamma3_minus_1_III(-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%rho0_III          (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%g_over_rho0_III   (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%p0_over_rho0_III  (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%rho0T0_III        (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%total_heat_III    (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%p1_III            (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%Xi_rsst_III       (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        allocate(B%v_wave_III        (-ng+1:ni+ng,-ng+1:nj+ng,-ng+1:nk+ng))
        ! 1-D profiles as ModBlock_InitSSM keeps them (3_6/ModBlock.f90:238-248)
        allocate(B%g_I(-ng+1:ni+ng),B%p0_I(-ng+1:ni+ng),B%rho0_I(-ng+1:ni+ng),B%te0_I(-ng+1:ni+ng))
        allocate(B%gamma1_I(-ng+1:ni+ng),B%gamma3_I(-ng+1:ni+ng),B%diffusion_I(-ng+1:ni+ng))
        allocate(B%cooling_I(-ng+1:ni+ng),B%Xi_rsst_I(-ng+1:ni+ng))
        B%g_I=bg(:,1); B%rho0_I=bg(:,2); B%p0_I=bg(:,3); B%te0_I=bg(:,4)
        B%gamma1_I=bg(:,5); B%gamma3_I=bg(:,6); B%diffusion_I=bg(:,7); B%cooling_I=0.0d0
        B%Xi_rsst_I=bg(:,8)
        B%if_SSM=.true.
        do j=-ng+1,nj+ng
            do k=-ng+1,nk+ng
                B%gamma1_III(:,j,k)        =bg(:,5)
                B%gamma3_minus_1_III(:,j,k)=bg(:,6)-1.0
                B%rho0_III(:,j,k)          =bg(:,2)
                B%g_over_rho0_III(:,j,k)   =bg(:,1)/bg(:,2)
                B%p0_over_rho0_III(:,j,k)  =bg(:,3)/bg(:,2)
                B%rho0T0_III(:,j,k)        =bg(:,2)*bg(:,4)
                B%total_heat_III(:,j,k)    =bg(:,7)+0.0d0
                B%Xi_rsst_III(:,j,k)       =bg(:,8)
            end do
        end do
    end subroutine harness_background

end program falcon_ref_harness
