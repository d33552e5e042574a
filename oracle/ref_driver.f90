! Harness for the reference kernels (test infrastructure only).
!
! Links the reference's own modules (integrals, twordms, onerdm,
! uniquemodule, compiled unmodified from their sources) and calls
!   set_P0, maxcoincidence, createtwordm, tot_integration,
!   set_P0_j2, tot_integration_j2, onerdm_creat, elastic_integration,
!   elastic_integration_j2, nuclei_electron_integration and the nuclear sinc
!   sum of types 5/6 (with calculate_form_factors' sinc),
!   unique_total
! on one input record.  The setup tables of variables_total /
! variables_elastic (Ddx/Ddy/Ddz, px/py/pz, e12) are read from the input
! record because src/Variables.f90 does not compile (line 1012); the z1/z2
! and elastic Z arrays are formed here exactly as src/Variables.f90:1041-1061
! and :1199-1216 form them.
!
! Usage: ref_driver in.bin out.bin   (stream-access, native endian)
program ref_driver
    use calculate_form_factors
    use integrals
    use twordms
    use onerdm
    use uniquemodule
    implicit none
    integer, parameter :: dp = selected_real_kind(15)
    integer, parameter :: ik = selected_int_kind(8)
    integer(ik) :: ngto, ng, nq, maxl, nmo, nconf, lconfs, nstates, state1, state2
    integer(ik) :: nmat, i, j, ii, jj, cc, cc2, maxnmo, cnt, dolist
    real(dp) :: cutoffz, cutoffmd, cutoffcentre, temp
    integer(ik), allocatable :: l(:), m(:), n(:), group(:), gs(:), gc(:)
    real(dp), allocatable :: ga(:), xx(:), yy(:), zz(:), q(:)
    real(dp), allocatable :: mmod(:,:), civs(:,:), px(:,:), py(:,:), pz(:,:)
    integer(ik), allocatable :: confs(:,:), ep3(:,:), ndiff2(:,:), mat(:,:), rmat(:,:)
    real(dp), allocatable :: ddx(:,:,:,:,:), ddy(:,:,:,:,:), ddz(:,:,:,:,:), e12(:,:,:)
    real(dp), allocatable :: p0(:,:,:,:), total(:), rtot(:), z1(:,:,:), z2(:,:,:)
    real(dp), allocatable :: zt(:,:,:), res1(:), res2(:), res3(:), res4(:), orm(:,:), zel(:,:)
    real(dp), allocatable :: res5(:), res6(:), Zn(:), geom(:,:)
    integer(ik) :: natom, na, nb, kq
    real(dp) :: Rij
    real(dp), allocatable :: p0j2(:,:,:,:)
    character(len=512) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) ngto, ng, nq, maxl, nmo, nconf, lconfs, nstates, state1, state2, dolist
    read(10) cutoffz, cutoffmd, cutoffcentre
    allocate(l(ngto), m(ngto), n(ngto), group(ngto), ga(ngto), xx(ngto), yy(ngto), zz(ngto))
    read(10) l, m, n, group
    read(10) ga, xx, yy, zz
    allocate(mmod(nmo, ngto), q(nq), confs(nconf, lconfs), civs(nconf, nstates))
    read(10) mmod
    read(10) q
    read(10) confs
    read(10) civs
    allocate(px(ng,ng), py(ng,ng), pz(ng,ng))
    read(10) px, py, pz
    allocate(ddx(maxl*2+1, maxl+1, maxl+1, ng, ng), ddy(maxl*2+1, maxl+1, maxl+1, ng, ng), &
             ddz(maxl*2+1, maxl+1, maxl+1, ng, ng), e12(nq, ng, ng))
    read(10) ddx, ddy, ddz
    read(10) e12
    read(10) natom
    allocate(Zn(natom), geom(natom, 3))
    read(10) Zn
    read(10) geom
    close(10)

    ! group metadata, src/main_calculation.f90:85-93
    allocate(gs(ng), gc(ng))
    do i = 1, ng
        do j = 1, ngto
            if (group(j) == i) then
                gs(i) = j
                gc(i) = count(group == i)
                exit
            end if
        end do
    end do

    allocate(p0(nq, 4*maxval(l)+1, 4*maxval(l)+1, 4*maxval(l)+1))
    p0 = 0.0_dp
    call set_P0(p0, 4*maxval(l), q)

    call maxcoincidence(confs, ep3, ndiff2)
    call createtwordm(confs, civs, ndiff2, ep3, mat, total, state1, state2)
    nmat = size(total)

    allocate(res1(nq), res2(nq), res3(nq), res4(nq))
    res1 = 0.0_dp
    res3 = 0.0_dp
    res4 = 0.0_dp
    ! the P0_j2 table of types 7 and 8 (src/main_calculation.f90:389,785)
    allocate(p0j2(nq, 4*maxval(l)+1, 4*maxval(l)+1, 4*maxval(l)+1))
    p0j2 = 0.0_dp
    call set_P0_j2(p0j2, 4*maxval(l), q)
    if (dolist /= 0) then
        ! z11/z22 of variables_total (src/Variables.f90:1041-1061); the pair
        ! sort of :1000-1008 does not change these values.
        allocate(z1(nmat, ngto, ngto), z2(nmat, ngto, ngto), zt(1,1,1))
        do ii = 1, ngto
            do jj = 1, ngto
                z1(:, ii, jj) = total * (mmod(min(mat(:,1),mat(:,2)), ii) * mmod(max(mat(:,1),mat(:,2)), jj) &
                              + mmod(min(mat(:,1),mat(:,2)), jj) * mmod(max(mat(:,1),mat(:,2)), ii))
                z2(:, ii, jj) = mmod(min(mat(:,3),mat(:,4)), ii) * mmod(max(mat(:,3),mat(:,4)), jj) &
                              + mmod(min(mat(:,3),mat(:,4)), jj) * mmod(max(mat(:,3),mat(:,4)), ii)
            end do
        end do
        zt = 0.0_dp
        call tot_integration(ng, px, py, pz, l, m, n, p0, ddx, ddy, ddz, z1, z2, zt, zt, &
                             gs, gc, group, cutoffz, cutoffmd, cutoffcentre, q, e12, res1)
        ! total j2 (type 7)
        call tot_integration_j2(ng, nq, px, py, pz, l, m, n, p0j2, ddx, ddy, ddz, z1, z2, &
                                gs, gc, group, cutoffz, cutoffmd, cutoffcentre, q, e12, res3)
        deallocate(z1, z2)
    end if

    call onerdm_creat(confs, civs, orm, maxnmo, state1, state2)
    ! Z of variables_elastic, src/Variables.f90:1199-1216
    allocate(zel(ngto, ngto))
    do ii = 1, ngto
        do jj = 1, ngto
            temp = 0.0_dp
            do cc = 1, maxnmo
                do cc2 = 1, maxnmo
                    temp = temp + orm(cc,cc2)*mmod(cc,ii)*mmod(cc2,jj)
                    temp = temp + orm(cc,cc2)*mmod(cc,jj)*mmod(cc2,ii)
                end do
            end do
            zel(ii,jj) = temp/2.0_dp
        end do
    end do
    call elastic_integration(ng, px, py, pz, l, m, n, p0, ddx, ddy, ddz, zel, gs, gc, group, &
                             cutoffz, cutoffmd, cutoffcentre, q, e12, res2)
    ! elastic j2 (type 8, src/main_calculation.f90:780-816)
    call elastic_integration_j2(ng, px, py, pz, l, m, n, p0j2, ddx, ddy, ddz, zel, gs, gc, group, &
                                cutoffz, cutoffmd, cutoffcentre, q, e12, res4)

    ! electron scattering terms (types 5/6, src/main_calculation.f90:298-312)
    allocate(res5(nq), res6(nq))
    call nuclei_electron_integration(Zn, geom, ng, px, py, pz, l, m, n, p0, ddx, ddy, ddz, zel, &
                                     gs, gc, group, cutoffz, cutoffmd, cutoffcentre, q, e12, res5)
    res6 = 0.0_dp
    do na = 1, size(Zn)
        do nb = 1, size(Zn)
            do kq = 1, nq
                Rij = sqrt(sum((geom(na,:) - geom(nb,:))**2))
                res6(kq) = res6(kq) + Zn(na) * Zn(nb) * sinc(q(kq) * abs(Rij))
            end do
        end do
    end do

    call unique_total(mat, total, rmat, rtot)
    cnt = size(rtot)

    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) nmat, maxnmo, cnt
    write(11) mat
    write(11) total
    write(11) res1
    write(11) res2
    write(11) orm
    write(11) zel
    write(11) rmat
    write(11) rtot
    write(11) res3
    write(11) res4
    write(11) res5
    write(11) res6
    write(11) p0
    close(11)
end program ref_driver
