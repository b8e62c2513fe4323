/*
 * Library set-up. Replaces include/sif/core/system.h of sif v0.5.0
 * (sif_config_t :75-91, sif_system_init/sif_system_finalise). Init binds the
 * calling thread's HIP device, creates the library stream and device pools;
 * finalise releases them. The finder returns SIF_ERR_UNINITIALISED without it
 * (reference include/sif/finder/exodus_finder.h:287-288).
 */
#ifndef SIF_CORE_SYSTEM_H
#define SIF_CORE_SYSTEM_H

#include <stdbool.h>
#include <stddef.h>

#include "sif/core/macros.h"

SIF_EXTERN_C_BEGIN

typedef struct {
  bool skip_tuning;
} sif_fft_config_t;

typedef struct {
  uint32_t n_threads;
} sif_omp_config_t;

typedef struct {
  const char* key;
  const char* value;
} sif_setting_t;

typedef struct {
  sif_fft_config_t* fft_config;
  sif_omp_config_t* omp_config;
  uint8_t log_level;
  const sif_setting_t* settings;
  size_t n_settings;
} sif_config_t;

/* SIF_OK, or SIF_ERR_UNSUPPORTED when no HIP device is visible. NULL config
 * takes the defaults. A second call is a no-op. */
SIF_API SIF_NODISCARD int sif_system_init(const sif_config_t* config);
SIF_API void sif_system_finalise(void);
SIF_API int sif_system_is_initialised(void);

SIF_EXTERN_C_END

#endif
