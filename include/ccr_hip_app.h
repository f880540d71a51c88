/* ccr_hip_app.h -- the Spec/Value front door on the HIP engine.
 *
 * The reference's front end calls its library through one entry point,
 * `Value ccr::app::call(const std::string& function, const Value& spec)` with a JSON-like
 * specification (market, models, portfolio, CSA, simulation settings) and a JSON-like result
 * (include/ccr/app/value.hpp:148-152; the entry points src/app/analyses.cpp:2376-2409). Errors
 * come back as { "error": "message" } rather than as exceptions (analyses.cpp:68-81).
 *
 * Here the specification and the result are JSON text. The spec keys, defaults and analyses are
 * the reference's (parseEnv / parseTrade / parseCsa, analyses.cpp:85-290; analysisJobs :682-780;
 * combine :1304-1447; the portfolio book :1604-2048), run on this library's engine instead of
 * the reference's fused kernels.
 */
#ifndef CCR_HIP_APP_H
#define CCR_HIP_APP_H

#include "ccr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Runs entry point `function` on the specification `spec_json` (a JSON object) and returns the
 * result as a NUL-terminated JSON text in *result_json, allocated by the library and released
 * with ccr_hip_app_free. Entry points (ccr_hip_app_functions):
 *   "version"     library and ABI version;
 *   "hipJobs"     the jobs of spec.analysis as the engine would compile them (parsed trades,
 *                 bootstrapped credit curves, grid sizes); needs no device (cf. gpuJobs,
 *                 analyses.cpp:2058-2095);
 *   "hipAnalyse"  runs spec.analysis -- "exposure", "cva", "xva", "collateral", "wrongWayRisk"
 *                 or "allocation", or a portfolio book (spec.scope = "book") of any of them --
 *                 on the HIP engine of `ctx` (cf. cpuKernels / gpuAnalyse,
 *                 analyses.cpp:2179-2214).
 * A failure of the analysis itself (a bad spec, an unknown analysis or entry point, a setup the
 * kernels do not implement) is a result { "error": "..." } (or { "unsupported": "..." }) with
 * CCR_OK, as the reference returns it. CCR_ERR_INVALID_ARGUMENT: a null function, spec or
 * result pointer, or no context for "hipAnalyse"; CCR_ERR_OUT_OF_MEMORY: no room for the
 * result. */
int32_t ccr_hip_app_call(ccr_hip_ctx* ctx, const char* function, const char* spec_json, char** result_json);

/* Releases a result of ccr_hip_app_call (NULL is ignored). */
void ccr_hip_app_free(char* result_json);

/* The entry points ccr_hip_app_call accepts, comma-separated (cf. ccr::app::functions,
 * value.hpp:151-152). */
const char* ccr_hip_app_functions(void);

#ifdef __cplusplus
}
#endif

#endif /* CCR_HIP_APP_H */
