// Build identity of the native-problem SDK.  A plugin records the value it was
// compiled with and the loader in libmoea_hip.so accepts it only when it
// equals its own: the plugin exchanges C++ objects (std::string,
// std::vector, std::variant, virtual calls) with the library, so both sides
// must come from the same SDK revision and C++ runtime.
#pragma once

#define CUDA_MOEA_SDK_STR2(x) #x
#define CUDA_MOEA_SDK_STR(x) CUDA_MOEA_SDK_STR2(x)

#if defined(_GLIBCXX_RELEASE)
#define CUDA_MOEA_SDK_CXXRT "libstdc++" CUDA_MOEA_SDK_STR(_GLIBCXX_RELEASE)
#elif defined(_LIBCPP_VERSION)
#define CUDA_MOEA_SDK_CXXRT "libc++" CUDA_MOEA_SDK_STR(_LIBCPP_VERSION)
#else
#define CUDA_MOEA_SDK_CXXRT "unknown-cxx-runtime"
#endif

#define CUDA_MOEA_NATIVE_BUILD_ID "moea_hip-sdk-1/" CUDA_MOEA_SDK_CXXRT "/gfx950"
