/* Error codes of the ufsecp C ABI.
 *
 * The numeric values are part of the ABI and are identical to the reference's
 * include/ufsecp/ufsecp_error.h:24-37 (CPU codes 0-13) and
 * include/ufsecp/ufsecp_gpu.h:80-86 (GPU codes 100-106), so callers compiled
 * against the reference keep their error handling unchanged. */
#ifndef UFSECP_ERROR_H
#define UFSECP_ERROR_H

typedef int ufsecp_error_t;

#define UFSECP_OK                  0
#define UFSECP_ERR_NULL_ARG        1
#define UFSECP_ERR_BAD_KEY         2
#define UFSECP_ERR_BAD_PUBKEY      3
#define UFSECP_ERR_BAD_SIG         4
#define UFSECP_ERR_BAD_INPUT       5
#define UFSECP_ERR_VERIFY_FAIL     6
#define UFSECP_ERR_ARITH           7
#define UFSECP_ERR_SELFTEST        8
#define UFSECP_ERR_INTERNAL        9
#define UFSECP_ERR_BUF_TOO_SMALL   10
#define UFSECP_ERR_NOT_FOUND       11
#define UFSECP_ERR_DEPRECATED_API  12
#define UFSECP_ERR_CANCELLED       13

#define UFSECP_ERR_GPU_UNAVAILABLE 100
#define UFSECP_ERR_GPU_DEVICE      101
#define UFSECP_ERR_GPU_LAUNCH      102
#define UFSECP_ERR_GPU_MEMORY      103
#define UFSECP_ERR_GPU_UNSUPPORTED 104
#define UFSECP_ERR_GPU_BACKEND     105
#define UFSECP_ERR_GPU_QUEUE       106

#ifndef UFSECP_API
#  define UFSECP_API __attribute__((visibility("default")))
#endif

#endif /* UFSECP_ERROR_H */
