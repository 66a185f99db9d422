/* Qualifier macros shared by the public headers.
 *
 * The public API is plain C.  The same headers are also read by HIP shader
 * modules (C++ compiled by hipcc), which need C linkage for the library entry
 * points and host+device definitions of the small math helpers that shaders
 * call. */
#pragma once

#ifdef __cplusplus
#  define SRP_EXTERN_C_BEGIN extern "C" {
#  define SRP_EXTERN_C_END }
#  define SRP_RESTRICT __restrict__
#else
#  define SRP_EXTERN_C_BEGIN
#  define SRP_EXTERN_C_END
#  define SRP_RESTRICT restrict
#endif

#if defined(__HIPCC__)
#  define SRP_HD __host__ __device__
#else
#  define SRP_HD
#endif

/* Shader modules compiled by hipcc get header-only host+device math.  The
 * library's own C translation unit that exports the math symbols defines
 * SRP_NO_INLINE_MATH. */
#if defined(__HIPCC__) && !defined(SRP_NO_INLINE_MATH)
#  define SRP_INLINE_MATH 1
#else
#  define SRP_INLINE_MATH 0
#endif
