! Harness for the reference's bit-string builders (test infrastructure only).
!
! Calls mcci_to_bit, createtwordm_bit and one_rdm_bit of the reference's
! module onerdm (src/one_rdm.f90, compiled unmodified) on the text file
! in.dat of the current directory (lines "i c1 c2 alpha beta", the layout
! the three routines read).
!
! Usage: ref_bit MODE n out.bin
!   conv : mcci_to_bit('in.dat', 'cv.dat'); cv.dat read back with
!          createtwordm_bit's own read statement -> c1(n), c2(n)
!   two  : createtwordm_bit('in.dat') -> nmat, mat(nmat,4), total(nmat)
!   route: mcci_to_bit('in.dat', 'cv.dat') then createtwordm_bit('cv.dat'),
!          the type-1 route of total_scattering_calculation_2
!   one  : one_rdm_bit('in.dat') -> maxnmo, onerdm(maxnmo,maxnmo)
!   fci  : createtwordm_bit_fci('in.dat', ...) with newdat, irep, start and
!          end read from fci.bin (stream: nuniq, newdat(nuniq,2), irep,
!          start, end, all int64) -> nmat, mat(nmat,4), total(nmat)
program ref_bit
    use onerdm
    implicit none
    integer, parameter :: dp = selected_real_kind(15)
    integer, parameter :: ik = selected_int_kind(8)
    character(len=16) :: mode, sn
    character(len=512) :: fout
    character(len=30) :: f30
    character(len=60) :: f60a, f60b
    integer(ik) :: n, i, maxnmo, nmat
    integer :: n4, dummy
    integer(ik), allocatable :: mat(:,:)
    real(dp), allocatable :: total(:), g(:,:), c1(:), c2(:)
    integer*8, allocatable :: newdat(:,:), irep(:), st(:), en(:)
    integer*8 :: nu
    character(len=20) :: f20
    integer :: a4, b4

    call get_command_argument(1, mode)
    call get_command_argument(2, sn)
    call get_command_argument(3, fout)
    read(sn, *) n
    n4 = n
    open(20, file=trim(fout), access='stream', form='unformatted', status='replace')
    select case (trim(mode))
    case ('conv')
        f60a = 'in.dat'
        f60b = 'cv.dat'
        call mcci_to_bit(f60a, f60b, n)
        allocate(c1(n), c2(n))
        open(15, file='cv.dat')
        do i = 1, n
            read(15, *) dummy, c1(i), c2(i), a4, b4
        end do
        close(15)
        write(20) c1, c2
    case ('two', 'route')
        if (trim(mode) == 'route') then
            f60a = 'in.dat'
            f60b = 'cv.dat'
            call mcci_to_bit(f60a, f60b, n)
            f30 = 'cv.dat'
        else
            f30 = 'in.dat'
        end if
        call createtwordm_bit(f30, n4, mat, total)
        nmat = size(total)
        write(20) nmat
        write(20) mat(:, 1), mat(:, 2), mat(:, 3), mat(:, 4)
        write(20) total
    case ('one')
        f30 = 'in.dat'
        allocate(newdat(1, 1), irep(1))
        newdat = 0
        irep = 0
        maxnmo = 0
        call one_rdm_bit(f30, g, maxnmo, n, newdat, irep)
        write(20) maxnmo
        write(20) g
    case ('fci')
        open(11, file='fci.bin', access='stream', form='unformatted', status='old')
        read(11) nu
        allocate(newdat(nu, 2), irep(nu), st(nu), en(nu))
        read(11) newdat, irep, st, en
        close(11)
        f20 = 'in.dat'
        call createtwordm_bit_fci(f20, n, newdat, irep, st, en, mat, total)
        nmat = size(total)
        write(20) nmat
        write(20) mat(:, 1), mat(:, 2), mat(:, 3), mat(:, 4)
        write(20) total
    end select
    close(20)
end program ref_bit
