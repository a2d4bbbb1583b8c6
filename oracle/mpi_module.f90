! The MPI module that the reference's `use MPI` names, compiled by this
! toolchain's flang from MPICH's own Fortran header.  The image's MPICH ships
! its mpi.mod in gfortran's format, which flang cannot read; the header holds the
! same constants and the library (libmpifort, libmpi) the same bindings.
! Test infrastructure only (oracle/Makefile, reference build).
module MPI
    include 'mpif.h'
end module MPI
