!***************************************************************
!> TEST HARNESS ONLY. Minimal modules named KindType and Domain that carry
!> exactly the entities the MODULE ADS shim (mod_ADS_hip.F90) reads, so the
!> shim can be compiled and exercised here without the rest of FALL3D. In a
!> FALL3D build the shim compiles against FALL3D's own modules instead.
!> Entity names and meaning follow Sources/mod_KindType.F90:15-22,42,465-502
!> and the serial decomposition of Sources/mod_Domain.F90:234-330.
!***************************************************************
MODULE KindType
  implicit none
  save
  integer,     parameter :: ip = 4
#if defined WITH_R4
  integer,     parameter :: rp = 4
#else
  integer,     parameter :: rp = 8
#endif
  real(rp),    parameter :: EPSILON = 1e-12_rp
  !
  type ARAKAWA_C_GRID
     real(rp), allocatable :: dX1_p(:)   ! (my_ips_2h:my_ipe_2h)
     real(rp), allocatable :: dX1_b(:)   ! (my_ibs_1h:my_ibe_1h)
     real(rp), allocatable :: dX2_p(:)
     real(rp), allocatable :: dX2_b(:)
     real(rp), allocatable :: dX3_p(:)
     real(rp), allocatable :: dX3_b(:)
  end type ARAKAWA_C_GRID
END MODULE KindType

MODULE Domain
  use KindType
  implicit none
  save
  integer(ip) :: my_ips, my_ipe, my_jps, my_jpe, my_kps, my_kpe
  integer(ip) :: my_ibs, my_ibe, my_jbs, my_jbe, my_kbs, my_kbe
  integer(ip) :: my_ips_2h, my_ipe_2h, my_jps_2h, my_jpe_2h, my_kps_2h, my_kpe_2h
  integer(ip) :: my_ibs_1h, my_ibe_1h, my_jbs_1h, my_jbe_1h, my_kbs_1h, my_kbe_1h
  integer(ip) :: my_W_proc = -1, my_E_proc = -1, my_S_proc = -1, my_N_proc = -1, my_U_proc = -1, my_D_proc = -1
CONTAINS
  !> serial decomposition (every neighbour is a physical boundary)
  subroutine domain_serial(nx, ny, nz)
    integer(ip), intent(IN) :: nx, ny, nz
    my_ips = 1; my_ipe = nx; my_jps = 1; my_jpe = ny; my_kps = 1; my_kpe = nz
    my_ibs = my_ips; my_ibe = my_ipe + 1
    my_jbs = my_jps; my_jbe = my_jpe + 1
    my_kbs = my_kps; my_kbe = my_kpe + 1
    my_ips_2h = my_ips - 2; my_ipe_2h = my_ipe + 2
    my_jps_2h = my_jps - 2; my_jpe_2h = my_jpe + 2
    my_kps_2h = my_kps - 2; my_kpe_2h = my_kpe + 2
    my_ibs_1h = my_ibs - 1; my_ibe_1h = my_ibe + 1
    my_jbs_1h = my_jbs - 1; my_jbe_1h = my_jbe + 1
    my_kbs_1h = my_kbs - 1; my_kbe_1h = my_kbe + 1
  end subroutine domain_serial
END MODULE Domain
