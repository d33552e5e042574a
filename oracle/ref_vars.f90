! Harness for the reference's variables_total (test infrastructure only).
!
! Links the reference's module `variables` (src/Variables.f90, compiled
! from a copy under _ref/ in which the one token that the compilers reject
! at line 1012, `divided==.True.`, reads `divided.eqv..True.`; nothing else
! differs) and calls variables_total on one input record: fill_md_table's
! repacked Hermite tables ddx/ddy/ddz, the pair centres px/py/pz, e12 (with
! the reference's single-precision pi) and the z11/z22 2-RDM factors.
!
! Usage: ref_vars in.bin out.bin   (stream access, native endian)
! in:  int32 ngto, ng, nq, maxl, nmo, nmat; int32 l, m, n, group [ngto];
!      f64 ga, xx, yy, zz [ngto]; f64 mmod (nmo, ngto); int32 mat (nmat, 4);
!      f64 total [nmat]; f64 q [nq]
! out: f64 px, py, pz (ng, ng); ddx, ddy, ddz (2maxl+1, maxl+1, maxl+1, ng,
!      ng); e12 (nq, ng, ng); z11, z22 (nmat, ngto, ngto)
program ref_vars
    use variables
    implicit none
    integer, parameter :: dp = selected_real_kind(15)
    integer, parameter :: ik = selected_int_kind(8)
    integer(ik) :: ngto, ng, nq, maxl, nmo, nmat, i, j
    integer(ik), allocatable :: l(:), m(:), n(:), group(:), gs(:), gc(:), mat(:,:)
    real(dp), allocatable :: ga(:), xx(:), yy(:), zz(:), q(:), total(:), mmod(:,:)
    real(dp), allocatable :: px(:,:), py(:,:), pz(:,:), e12(:,:,:)
    real(dp), allocatable :: ddx(:,:,:,:,:), ddy(:,:,:,:,:), ddz(:,:,:,:,:)
    real(dp), allocatable :: z11(:,:,:), z22(:,:,:), z1t(:,:,:), z2t(:,:,:)
    character(len=512) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) ngto, ng, nq, maxl, nmo, nmat
    allocate(l(ngto), m(ngto), n(ngto), group(ngto), ga(ngto), xx(ngto), yy(ngto), zz(ngto))
    read(10) l, m, n, group
    read(10) ga, xx, yy, zz
    allocate(mmod(nmo, ngto), mat(nmat, 4), total(nmat), q(nq))
    read(10) mmod
    read(10) mat
    read(10) total
    read(10) q
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
    allocate(px(ng,ng), py(ng,ng), pz(ng,ng), e12(nq,ng,ng))
    allocate(ddx(maxl*2+1, maxl+1, maxl+1, ng, ng), ddy(maxl*2+1, maxl+1, maxl+1, ng, ng), &
             ddz(maxl*2+1, maxl+1, maxl+1, ng, ng))
    ddx = 0.0_dp
    ddy = 0.0_dp
    ddz = 0.0_dp
    call variables_total(px, py, pz, ddx, ddy, ddz, z11, z22, z1t, z2t, e12, maxl, ngto, ng, &
                         gs, gc, group, ga, l, m, n, xx, yy, zz, mmod, mat(:,1), mat(:,2), &
                         mat(:,3), mat(:,4), nmat, total, q, nq)
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) px, py, pz
    write(11) ddx, ddy, ddz
    write(11) e12
    write(11) z11, z22
    close(11)
end program ref_vars
