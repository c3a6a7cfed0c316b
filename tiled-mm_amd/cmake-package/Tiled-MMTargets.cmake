# Imported target Tiled-MM::Tiled-MM. This file lies in <libdir>/cmake/Tiled-MM, both where
# `make -C tiled-mm_amd` builds the library (tiled-mm_amd/lib, headers in the repository's
# include/) and where `make -C tiled-mm_amd install PREFIX=...` puts it (PREFIX/lib, headers in
# PREFIX/include).
get_filename_component(_tmm_libdir "${CMAKE_CURRENT_LIST_DIR}/../.." ABSOLUTE)
get_filename_component(_tmm_prefix "${_tmm_libdir}/.." ABSOLUTE)
if(EXISTS "${_tmm_prefix}/include/tiledmm.h")
  set(_tmm_incdir "${_tmm_prefix}/include")
else()
  get_filename_component(_tmm_incdir "${_tmm_prefix}/../include" ABSOLUTE)
endif()
if(NOT EXISTS "${_tmm_libdir}/libtiledmm.so" OR NOT EXISTS "${_tmm_incdir}/Tiled-MM/tiled_mm.hpp")
  set(Tiled-MM_FOUND FALSE)
  set(Tiled-MM_NOT_FOUND_MESSAGE "libtiledmm.so or the Tiled-MM headers are missing next to ${CMAKE_CURRENT_LIST_DIR}")
  return()
endif()

set(_tmm_rocm "$ENV{ROCM_PATH}")
if(NOT _tmm_rocm)
  set(_tmm_rocm "/opt/rocm")
endif()
find_library(TILEDMM_AMDHIP64 amdhip64 HINTS "${_tmm_rocm}/lib" NO_DEFAULT_PATH)
find_library(TILEDMM_AMDHIP64 amdhip64)
find_path(TILEDMM_HIP_INCLUDE hip/hip_runtime_api.h HINTS "${_tmm_rocm}/include" NO_DEFAULT_PATH)
find_path(TILEDMM_HIP_INCLUDE hip/hip_runtime_api.h)
if(NOT TILEDMM_AMDHIP64 OR NOT TILEDMM_HIP_INCLUDE)
  set(Tiled-MM_FOUND FALSE)
  set(Tiled-MM_NOT_FOUND_MESSAGE "the HIP runtime (libamdhip64, hip/hip_runtime_api.h) was not found under ${_tmm_rocm}")
  return()
endif()

add_library(Tiled-MM::Tiled-MM SHARED IMPORTED)
set_target_properties(Tiled-MM::Tiled-MM PROPERTIES
  IMPORTED_LOCATION "${_tmm_libdir}/libtiledmm.so"
  IMPORTED_SONAME "libtiledmm.so"
  INTERFACE_INCLUDE_DIRECTORIES "${_tmm_incdir};${TILEDMM_HIP_INCLUDE}"
  INTERFACE_COMPILE_DEFINITIONS "__HIP_PLATFORM_AMD__"
  INTERFACE_COMPILE_FEATURES "cxx_std_14"
  INTERFACE_LINK_LIBRARIES "${TILEDMM_AMDHIP64}")
unset(_tmm_libdir)
unset(_tmm_prefix)
unset(_tmm_incdir)
unset(_tmm_rocm)
