// host_sanitize_stubs.cpp -- the symbols the planners' translation units
// reference from the device side of the library (kernel launchers, the
// profiler, the HIP runtime), for the sanitizer build of the host code
// (tests/test_host_sanitizers.py).  The planner entry points never reach them;
// each stub reports itself and fails.  The error bookkeeping (fail,
// qsv_last_error) is the library's own behaviour restated.
#include <cstdio>
#include <string>
#include <vector>

#include "../../qpanda_amd/csrc/qsv_fusion.h"
#include "../../qpanda_amd/csrc/qsv_sched.h"

namespace {
thread_local std::string g_err;
int unreachable(const char *what)
{
    std::fprintf(stderr, "stub reached: %s\n", what);
    return QSV_ERR_UNSUPPORTED;
}
}  // namespace

namespace qsv {
void set_error(const std::string &m) { g_err = m; }
int fail(int code, const std::string &m)
{
    g_err = m;
    return code;
}
int hip_fail(hipError_t, const char *what) { return fail(QSV_ERR_DEVICE, what); }
int Profiler::begin() { return unreachable("Profiler::begin"); }
int Profiler::end(int, double, double, int64_t) { return unreachable("Profiler::end"); }
int f16_state_scale(qsv_state *) { return unreachable("f16_state_scale"); }
size_t fused_operand_floats(int, int) { return 0; }
void fused_operands(int, const double *, int, float *) { unreachable("fused_operands"); }
int launch_block(qsv_state *, const int *, int, const double *, uint64_t) { return unreachable("launch_block"); }
int launch_fused_dev(qsv_state *, const int *, int, const float *, int, int) { return unreachable("launch_fused_dev"); }
int launch_pass_dev(qsv_state *, const std::vector<const std::vector<int> *> &, const std::vector<const float *> &, int,
                    int, int, const std::vector<std::pair<int, int>> *)
{
    return unreachable("launch_pass_dev");
}
int pass_tile_bits(const std::vector<const std::vector<int> *> &, int) { return -1; }
void launch_features(int, int, const std::vector<const std::vector<int> *> &, bool, int, float *)
{
    unreachable("launch_features");
}
int group_launches(qsv_state *, const std::vector<PlanItem> &, const qsv_run_opts &, std::vector<Launch> &, double *,
                   const std::vector<int> *)
{
    return unreachable("group_launches");
}
int plan_learned(qsv_state *, const qsv_gate *, int64_t, const double *, const qsv_run_opts &, std::vector<PlanItem> &,
                 SchedChoice *, std::vector<int> *)
{
    return unreachable("plan_learned");
}
}  // namespace qsv

extern "C" {
const char *qsv_last_error(void) { return g_err.c_str(); }
int qsv_measure(qsv_state *, int, double, int *) { return unreachable("qsv_measure"); }
int qsv_x(qsv_state *, int, uint64_t) { return unreachable("qsv_x"); }
int32_t qsv_pm_next(int32_t s) { return s; }
int32_t qsv_pm_jump(int32_t s, uint64_t) { return s; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipFree(void *) { return hipErrorNotSupported; }
hipError_t hipMalloc(void **, size_t) { return hipErrorNotSupported; }
hipError_t hipMemcpyAsync(void *, const void *, size_t, hipMemcpyKind, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipErrorNotSupported; }
hipError_t hipEventCreateWithFlags(hipEvent_t *, unsigned) { return hipErrorNotSupported; }
hipError_t hipEventDestroy(hipEvent_t) { return hipErrorNotSupported; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipErrorNotSupported; }
hipError_t hipHostMalloc(void **, size_t, unsigned) { return hipErrorNotSupported; }
hipError_t hipHostFree(void *) { return hipErrorNotSupported; }
hipError_t hipSetDevice(int) { return hipErrorNotSupported; }
}
