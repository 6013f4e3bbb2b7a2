! backup_writer.f90 -- test infrastructure (part of the oracle): writes a
! particle backup file with the same Fortran statements as the reference's
! writer (main.f90:1125-1131), so the Fortran runtime of this image defines
! the bytes that the library's writer must reproduce and its reader accept.
!
! usage: backup_writer PATH NTYPE NPROC NP(1,1) NP(2,1) ... NP(NTYPE,NPROC)
! vxp(k,i,ptype,iproc) = k + 8*i + 1024*ptype + 65536*iproc + 0.125
program backup_writer
  implicit none
  integer :: ntype, nproc, ip, p, i, k, nmax, narg, gtype
  integer, allocatable :: np_tot(:,:)
  real(kind=8), allocatable :: vxp(:,:,:,:)
  real(kind=8) :: time, pabs_cor
  character(len=512) :: path, arg

  narg = command_argument_count()
  call get_command_argument(1, path)
  call get_command_argument(2, arg); read(arg,*) ntype
  call get_command_argument(3, arg); read(arg,*) nproc
  allocate(np_tot(ntype,nproc))
  do ip = 1, nproc
     do p = 1, ntype
        call get_command_argument(3 + p + ntype*(ip-1), arg)
        read(arg,*) np_tot(p,ip)
     enddo
  enddo
  nmax = max(1, maxval(np_tot))
  allocate(vxp(6,nmax,ntype,nproc))
  do ip = 1, nproc
     do p = 1, ntype
        do i = 1, nmax
           do k = 1, 6
              vxp(k,i,p,ip) = dble(k + 8*i + 1024*p + 65536*ip) + 0.125d0
           enddo
        enddo
     enddo
  enddo
  time = 1.5d-6
  gtype = 7
  pabs_cor = 0.75d0
  open(41, file=trim(path), form='UNFORMATTED')
  write(41) time,gtype,pabs_cor
  write(41) np_tot(1:ntype,1:nproc)
  do ip = 1, nproc
     do p = 1, ntype
        write(41) vxp(:,1:np_tot(p,ip),p,ip)
     enddo
  enddo
  close(41)
end program backup_writer
