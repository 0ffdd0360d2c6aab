/* spektralab.h -- the public C ABI of the MI355X render engine.
 *
 * Drop-in for the reference engine's include/spektralab.h: the same 26 entry points, the same
 * status codes and the same LP64 struct layouts (static-asserted below), so a host that was
 * written against the reference header links against libspektralab.so unchanged.
 *
 * Conventions of every entry point:
 *   - C linkage; only spk_* / spkx_* symbols are exported from the shared library.
 *   - failure: a negative spk_status, or a NULL handle, with a message in a per-thread
 *     last-error slot (spk_last_error).  Exceptions never cross this boundary.
 *   - inputs are borrowed for the duration of the call; the engine copies what it keeps.
 *   - out_json: a heap string owned by the caller, released with spk_string_free.
 *   - a returned const char*: engine-owned, never freed by the caller.
 *   - calls on different sessions may run concurrently; calls on one session are serialised
 *     by that session's lock; spk_cancel and spk_progress never take that lock.
 *
 * MI355X mapping of the device-typed fields:
 *   - `device` (engine creation): NULL, or a pointer to an int32_t HIP device ordinal.
 *   - spk_device_image.buffer: a device pointer (hipMalloc'd) on the engine's device.
 *   - spk_result.texture: an opaque +1 handle owning the rendered image (its device buffer and
 *     the pinned host copy that rgba16 points into); released by spk_result_free.
 */
#ifndef SPEKTRALAB_H
#define SPEKTRALAB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(SPK_BUILDING_LIBRARY)
#define SPK_API __attribute__((visibility("default")))
#else
#define SPK_API
#endif

/* Version numbers (also reported by spk_capabilities / spk_params_schema). */
#define SPK_TRANSPORT_VERSION 1
#define SPK_SCHEMA_VERSION 1
#define SPK_VERSION_MAJOR 1
#define SPK_VERSION_MINOR 0
#define SPK_VERSION_PATCH 0
#define SPK_VERSION_STRING "1.0.0"
#define SPK_ENGINE_NAME "spektralab.native"

/* Status codes; there are no others.
 *   SPK_ERR_INVALID_ARG  a required pointer is NULL, unknown tier, unparsable gamut JSON
 *   SPK_ERR_USER         unknown/invalid parameter, stock, solve target or progress id
 *   SPK_ERR_RESOURCE     reserved (never returned)
 *   SPK_ERR_UNSUPPORTED  reserved (never returned)
 *   SPK_ERR_GPU          device/launch failure, out of memory, any non-cancel render failure
 *   SPK_ERR_IO           resource file missing or corrupt
 *   SPK_ERR_CANCELLED    the render was cancelled
 *   SPK_ERR_INTERNAL     an engine bug
 */
typedef int32_t spk_status;
#define SPK_OK 0
#define SPK_ERR_INVALID_ARG (-1)
#define SPK_ERR_USER (-2)
#define SPK_ERR_RESOURCE (-3)
#define SPK_ERR_UNSUPPORTED (-4)
#define SPK_ERR_GPU (-5)
#define SPK_ERR_IO (-6)
#define SPK_ERR_CANCELLED (-7)
#define SPK_ERR_INTERNAL (-8)

/* Device types, fixed for an engine's life. */
#define SPK_DEVICE_DESKTOP 0
#define SPK_DEVICE_PHONE 1
#define SPK_DEVICE_TABLET 2

/* Largest frame accepted by spk_open: 14204 x 10652 pixels. */
#define SPK_MAX_PIXELS 151301008ULL

typedef struct spk_engine spk_engine;
typedef struct spk_session spk_session;

/* Host image, borrowed for the call.  24 bytes, 8-aligned.
 * Pixel (x, y) channel c is data[(y * width + x) * channels + c]; rows top to bottom;
 * channels 3 or 4 (a 4th channel is dropped); non-finite values become 0 on copy. */
typedef struct spk_image {
    const float* data;
    uint32_t width;
    uint32_t height;
    uint32_t channels;
} spk_image;

/* The same image already in device memory of the engine's GPU.  24 bytes, 8-aligned. */
typedef struct spk_device_image {
    void* buffer;
    uint32_t width;
    uint32_t height;
    uint32_t channels;
} spk_device_image;

/* One rendered image.  88 bytes, 8-aligned.
 * Zero-filled on entry by every render call (a failed call leaves it zeroed).
 * Pixel (x, y) channel c is rgba16[4 * (y * row_stride_px + x) + c]; alpha = 65535.
 * texture is +1: release it with spk_result_free.  The image outlives session and engine. */
typedef struct spk_result {
    const uint16_t* rgba16;
    void* texture;
    uint32_t width;
    uint32_t height;
    uint32_t row_stride_px;
    double elapsed_ms;
    int32_t reprint;
    int32_t negative_was_cached;
    char progress_id[40];
} spk_result;

#ifdef __cplusplus
#define SPK_STATIC_ASSERT(c, m) static_assert(c, m)
#define SPK_ALIGNOF(t) alignof(t)
#else
#define SPK_STATIC_ASSERT(c, m) _Static_assert(c, m)
#define SPK_ALIGNOF(t) _Alignof(t)
#endif

SPK_STATIC_ASSERT(sizeof(void*) == 8, "LP64 required");
SPK_STATIC_ASSERT(sizeof(spk_image) == 24 && SPK_ALIGNOF(spk_image) == 8, "spk_image layout");
SPK_STATIC_ASSERT(offsetof(spk_image, data) == 0 && offsetof(spk_image, width) == 8 &&
                      offsetof(spk_image, height) == 12 && offsetof(spk_image, channels) == 16,
                  "spk_image fields");
SPK_STATIC_ASSERT(sizeof(spk_device_image) == 24 && SPK_ALIGNOF(spk_device_image) == 8,
                  "spk_device_image layout");
SPK_STATIC_ASSERT(offsetof(spk_device_image, buffer) == 0 && offsetof(spk_device_image, width) == 8 &&
                      offsetof(spk_device_image, height) == 12 && offsetof(spk_device_image, channels) == 16,
                  "spk_device_image fields");
SPK_STATIC_ASSERT(sizeof(spk_result) == 88 && SPK_ALIGNOF(spk_result) == 8, "spk_result layout");
SPK_STATIC_ASSERT(offsetof(spk_result, rgba16) == 0 && offsetof(spk_result, texture) == 8 &&
                      offsetof(spk_result, width) == 16 && offsetof(spk_result, height) == 20 &&
                      offsetof(spk_result, row_stride_px) == 24 && offsetof(spk_result, elapsed_ms) == 32 &&
                      offsetof(spk_result, reprint) == 40 && offsetof(spk_result, negative_was_cached) == 44 &&
                      offsetof(spk_result, progress_id) == 48,
                  "spk_result fields");

/* ---- engine lifecycle ----------------------------------------------------------------------
 * Reference: include/spektralab.h:181-208, src/abi/abi.cpp:94-125. */

/* Load the resources folder (constants blob, profiles on demand, optional filter DB and
 * print-LUT index), select the HIP device, run the numeric probe, build the schema.
 * NULL on failure with the reason in spk_last_error(NULL). */
SPK_API spk_engine* spk_engine_create(const char* resources_dir, void* device);
SPK_API spk_engine* spk_engine_create_for_device(const char* resources_dir, void* device,
                                                 int32_t device_type);
/* Releases every open session, then the engine; results already handed out stay valid. */
SPK_API void spk_engine_destroy(spk_engine* engine);
/* Frees idle scratch device memory; returns the bytes freed. */
SPK_API size_t spk_engine_trim(spk_engine* engine);
/* "spektralab-native <semver> (<date>, math=safe)", static. */
SPK_API const char* spk_build_info(void);

/* ---- introspection (engine-owned JSON; NULL engine -> "") ----------------------------------
 * Reference: include/spektralab.h:210-224. */
SPK_API const char* spk_capabilities(spk_engine* engine);
SPK_API const char* spk_params_schema(spk_engine* engine);
SPK_API const char* spk_memory_report(spk_engine* engine);

/* ---- warm-up (include/spektralab.h:226-238) ------------------------------------------------ */
SPK_API spk_status spk_warm_up(spk_engine* engine, const char* film_stock, const char* print_stock,
                               char** out_json);

/* ---- sessions (include/spektralab.h:240-297) ------------------------------------------------
 * spk_open copies the frame into the session's device plane (non-finite -> 0), validates the
 * delta, resolves stocks and builds the float64 set-up of the film and print sides. */
SPK_API spk_session* spk_open(spk_engine* engine, const spk_image* input, const char* params_delta_json,
                              char** out_json);
SPK_API spk_session* spk_open_device(spk_engine* engine, const spk_device_image* input,
                                     const char* params_delta_json, char** out_json);
SPK_API spk_status spk_set_params(spk_session* session, const char* params_delta_json, char** out_json);
SPK_API spk_status spk_get_params(spk_session* session, char** out_json);
SPK_API void spk_session_release(spk_session* session);

/* ---- rendering (include/spektralab.h:299-324) -----------------------------------------------
 * tier = "live" | "preview" | "full".
 * spk_render develops afresh (new film seed when grain is on), caches the negative, prints.
 * spk_reprint reuses the tier's cached negative when there is one (reprint = 1). */
SPK_API spk_status spk_render(spk_session* session, const char* tier, spk_result* out);
SPK_API spk_status spk_reprint(spk_session* session, const char* tier, spk_result* out);
SPK_API void spk_result_free(spk_result* result);

/* ---- solve, print-preview tables, output transform (include/spektralab.h:326-371) ---------- */
SPK_API spk_status spk_solve(spk_session* session, const char* target, char** out_json);
SPK_API const char* spk_print_lut_catalog(spk_engine* engine);
SPK_API spk_status spk_print_lut_table(spk_engine* engine, const char* print_stock, const float** out_table,
                                       uint32_t* out_size);
SPK_API spk_status spk_preview_stock_lut(spk_session* session, const char* print_stock, const char* tier,
                                         spk_result* out, char** out_json);
SPK_API spk_status spk_output_transform(spk_engine* engine, const char* src_cs, const char* dst_cs,
                                        const char* gamut_compress_json, char** out_json,
                                        const float** out_cmax, uint32_t* out_cmax_count);

/* ---- control (include/spektralab.h:373-386) ------------------------------------------------- */
SPK_API void spk_cancel(spk_session* session, const char* progress_id);
SPK_API spk_status spk_progress(spk_session* session, const char* progress_id, char** out_json);
SPK_API void spk_string_free(char* s);
SPK_API const char* spk_last_error(spk_engine* engine);

#ifdef __cplusplus
}
#endif

#endif /* SPEKTRALAB_H */
