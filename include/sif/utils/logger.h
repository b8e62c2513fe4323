/*
 * The error record of the last failed call on this thread. Replaces
 * include/sif/utils/logger.h:297-349 of sif v0.5.0 (sif_error_message,
 * sif_error_status, sif_error_errno, sif_error_clear).
 */
#ifndef SIF_UTILS_LOGGER_H
#define SIF_UTILS_LOGGER_H

#include "sif/core/macros.h"

SIF_EXTERN_C_BEGIN

/* Message of the last error on this thread, "" if none. */
SIF_API const char* sif_error_message(void);
/* Status of the last error on this thread, SIF_OK if none. */
SIF_API int sif_error_status(void);
/* errno captured with the last SIF_ERR_IO, 0 otherwise. */
SIF_API int sif_error_errno(void);
SIF_API void sif_error_clear(void);

SIF_EXTERN_C_END

#endif
