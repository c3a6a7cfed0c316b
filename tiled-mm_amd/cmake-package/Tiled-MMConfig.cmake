# Package configuration of libtiledmm.so, the MI355X out-of-core GEMM library, under the package
# and target names of the reference (Tiled-MM CMakeLists.txt:44-63, cmake/Tiled-MMConfig.cmake.in):
#   find_package(Tiled-MM REQUIRED)
#   target_link_libraries(app PRIVATE Tiled-MM::Tiled-MM)
# gives the C++ drop-in headers (<Tiled-MM/tiled_mm.hpp>, ...), the C ABI (<tiledmm.h>) and the
# library. The headers use the HIP runtime API, found under ROCM_PATH (default /opt/rocm).
if(NOT TARGET Tiled-MM::Tiled-MM)
  include("${CMAKE_CURRENT_LIST_DIR}/Tiled-MMTargets.cmake")
endif()
set(Tiled-MM_LIBRARIES Tiled-MM::Tiled-MM)
