/*
 * sif-amd: MI355X-native exodus hot path with sif's C API.
 *
 * Scalar type, option bits and status codes shared by every public header.
 * The numeric values are the ones sif v0.5.0 uses, so a caller compiled
 * against the reference headers links against this library unchanged:
 *   sif_real, sif_option      include/sif/core/macros.h:42, :504-534
 *   SIF_FINDER_* option bits  include/sif/core/macros.h:73-248
 *   SIF_MESH_* option bits    include/sif/core/macros.h:269-293
 *   SIF_PBC_* option bits     include/sif/core/macros.h:52-54
 *   SIF_PROFILES_* bits       include/sif/core/macros.h:310-332
 *   SIF_OK ... SIF_ERR_*      include/sif/core/macros.h:690-707
 */
#ifndef SIF_CORE_MACROS_H
#define SIF_CORE_MACROS_H

#include <float.h>
#include <math.h>
#include <stdint.h>

#ifdef __cplusplus
#  define SIF_EXTERN_C_BEGIN extern "C" {
#  define SIF_EXTERN_C_END   }
#else
#  define SIF_EXTERN_C_BEGIN
#  define SIF_EXTERN_C_END
#endif

#if defined(__GNUC__) || defined(__clang__)
#  define SIF_NODISCARD __attribute__((warn_unused_result))
#  define SIF_API       __attribute__((visibility("default")))
#else
#  define SIF_NODISCARD
#  define SIF_API
#endif

/* Positions, grid values and radii are single precision, as in the
 * reference's default build. */
typedef float sif_real;

/* Bit set of SIF_* options. */
typedef uint32_t sif_option;

#define SIF_DEFAULT 0u

/* pi rounded to sif_real, the value every float formula of the path uses. */
#define SIF_PI ((sif_real)3.14159265358979323846)

/* --- finder options --- */

/* Search sphere radius as a multiple of the rung (bits 12-13). */
#define SIF_FINDER_SEARCH_1_50  (0u << 12)
#define SIF_FINDER_SEARCH_1_25  (1u << 12)
#define SIF_FINDER_SEARCH_1_75  (2u << 12)
#define SIF_FINDER_SEARCH_2_00  (3u << 12)
#define SIF__FINDER_SEARCH_MASK (3u << 12)

static inline float sif__finder_search_factor(uint32_t opt) {
  /* Indexed by the encoding; the default is the zero value. */
  static const float factors[4] = {1.5f, 1.25f, 1.75f, 2.0f};
  return factors[(opt & SIF__FINDER_SEARCH_MASK) >> 12];
}

/* Divide the assignment window as W/S (aliasing-aware) instead of 1/W. */
#define SIF_FINDER_ALIASING_WINDOW (1u << 8)
/* Candidate seeds: every cell under the threshold, or the local minima. */
#define SIF_FINDER_SEEDS_CELLS  (0u << 9)
#define SIF_FINDER_SEEDS_MINIMA (1u << 9)
#define SIF__FINDER_SEEDS_MASK  (1u << 9)
/* Write no checkpoint. */
#define SIF_FINDER_NO_CHECKPOINT (1u << 11)
/* Leave the assignment window in the spectrum. */
#define SIF_FINDER_KEEP_CIC_WINDOW (1u << 14)
/* Do not restore the grid after the run. */
#define SIF_FINDER_CONSUME_GRID (1u << 15)
/* Test overlap on the Lagrangian radius R (1 + threshold)^(1/3). */
#define SIF_FINDER_LAGRANGIAN_OVERLAP (1u << 18)
/* Where candidates sit (bits 19-20). */
#define SIF_FINDER_REFINE_NONE     (0u << 19)
#define SIF_FINDER_REFINE_GAUSSIAN (1u << 19)
#define SIF_FINDER_REFINE_TRACERS  (2u << 19)
#define SIF_FINDER_REFINE_HILL     (3u << 19)
#define SIF__FINDER_REFINE_MASK    (3u << 19)
/* Acceptance rules of the excursion-peak variant. */
#define SIF_FINDER_EXTEND_REACH (1u << 21)
#define SIF_FINDER_DISPLACE     (1u << 22)

/* --- chain mesh options --- */
#define SIF_MESH_DROP_INDICES (1u << 16)
#define SIF_MESH_NO_CANONICAL (1u << 17)

/* --- boundary (bit 0): a periodic box, or an open one with nothing outside --- */
#define SIF_PBC_PERIODIC (0u << 0)
#define SIF_PBC_OPEN     (1u << 0)
#define SIF__PBC_MASK    (1u << 0)

/* --- profile options --- */
/* A density bin holds the contrast enclosed within its outer edge, or that
 * of its own shell (bit 8). */
#define SIF_PROFILES_CUMULATIVE   (0u << 8)
#define SIF_PROFILES_DIFFERENTIAL (1u << 8)
#define SIF__PROFILES_BIN_MASK    (1u << 8)
/* A shell's radial velocity averaged over its tracers, or over their
 * weights (bit 9). */
#define SIF_PROFILES_VELOCITY_NUMBER   (0u << 9)
#define SIF_PROFILES_VELOCITY_WEIGHTED (1u << 9)
#define SIF__PROFILES_VELOCITY_MASK    (1u << 9)

/* --- status codes --- */
#define SIF_OK                0
#define SIF_ERR_ALLOC         (-1)
#define SIF_ERR_INVALID       (-2)
#define SIF_ERR_RANGE         (-3)
#define SIF_ERR_IO            (-4)
#define SIF_ERR_UNSUPPORTED   (-5)
#define SIF_ERR_UNINITIALISED (-6)

/* --- log levels --- */
#define SIF_LOG_LEVEL_TRACE   0
#define SIF_LOG_LEVEL_DEBUG   1
#define SIF_LOG_LEVEL_INFO    2
#define SIF_LOG_LEVEL_WARNING 3
#define SIF_LOG_LEVEL_ERROR   4
#define SIF_LOG_LEVEL_NONE    5

/* What a position triple holds. */
typedef enum {
  SIF_COORDINATES_CARTESIAN = 0,
  SIF_COORDINATES_SKY
} sif_coordinates_t;

#endif /* SIF_CORE_MACROS_H */
