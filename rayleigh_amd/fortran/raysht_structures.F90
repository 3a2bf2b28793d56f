!
! Container types of the transform interface, for building the RaySHT Fortran
! binding outside the Rayleigh tree (its tests).  Inside Rayleigh the tree's own
! Structures module (src/Data_Structures/Structures.F90:25-43) provides these types
! and this file is not compiled; the two declarations must stay layout-identical.
!
Module Structures
    Implicit None
    Type, Public :: rmcontainer
        Real*8, Allocatable :: data(:,:)
    End Type rmcontainer

    Type, Public :: rmcontainer3d
        Real*8, Allocatable :: data(:,:,:)
    End Type rmcontainer3d

    Type, Public :: rmcontainer4d
        Real*8, Allocatable :: data(:,:,:,:)
    End Type rmcontainer4d
End Module Structures
