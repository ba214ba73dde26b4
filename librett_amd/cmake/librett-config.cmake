# CMake package of the gfx950 transpose library: find_package(librett)
# gives the imported target librett::librett (the name the reference exports,
# cmake/librett-config.cmake.in and src/CMakeLists.txt:78-92), with
# LIBRETT_USES_HIP defined and the header directory on the include path
# (both <librett/librett.h> and <librett.h> resolve).
#
# Installed by `make -C librett_amd install PREFIX=...` as
#   PREFIX/include/librett/{librett.h,librett_c.h}
#   PREFIX/lib/{librett.so,selector_weights.bin}
#   PREFIX/lib/cmake/librett/{librett-config.cmake,librett-config-version.cmake}
get_filename_component(_librett_prefix "${CMAKE_CURRENT_LIST_DIR}/../../.." ABSOLUTE)

if(NOT TARGET librett::librett)
  add_library(librett::librett SHARED IMPORTED)
  set_target_properties(librett::librett PROPERTIES
    IMPORTED_LOCATION "${_librett_prefix}/lib/librett.so"
    IMPORTED_SONAME "librett.so"
    INTERFACE_INCLUDE_DIRECTORIES "${_librett_prefix}/include;${_librett_prefix}/include/librett"
    INTERFACE_COMPILE_DEFINITIONS "LIBRETT_USES_HIP")
  # The header includes <hip/hip_runtime.h>.
  if(NOT DEFINED ROCM_PATH)
    if(DEFINED ENV{ROCM_PATH})
      set(ROCM_PATH "$ENV{ROCM_PATH}")
    else()
      set(ROCM_PATH "/opt/rocm")
    endif()
  endif()
  find_package(hip QUIET HINTS "${ROCM_PATH}" "${ROCM_PATH}/lib/cmake/hip")
  if(TARGET hip::host)
    set_property(TARGET librett::librett APPEND PROPERTY INTERFACE_LINK_LIBRARIES hip::host)
  else()
    set_property(TARGET librett::librett APPEND PROPERTY INTERFACE_INCLUDE_DIRECTORIES "${ROCM_PATH}/include")
    set_property(TARGET librett::librett APPEND PROPERTY INTERFACE_COMPILE_DEFINITIONS "__HIP_PLATFORM_AMD__")
  endif()
endif()

set(LIBRETT_FOUND TRUE)
set(LIBRETT_LIBRARIES librett::librett)
set(LIBRETT_INCLUDE_DIRS "${_librett_prefix}/include")
unset(_librett_prefix)
