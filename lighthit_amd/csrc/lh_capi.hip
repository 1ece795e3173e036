// Error reporting and version of the lighthit_amd C-ABI.
#include "lh_common.h"

#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace lh {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, args);
    va_end(args);
}

namespace {

struct Mark {
    std::string name;
    hipEvent_t start;
    hipEvent_t end;
};

std::atomic<bool> g_timing{false};
std::mutex g_marks_lock;
std::vector<Mark> g_marks;

}  // namespace

bool timing_enabled() { return g_timing.load(std::memory_order_relaxed); }

hipEvent_t timing_start(hipStream_t stream) {
    hipEvent_t start = nullptr;
    if (hipEventCreate(&start) != hipSuccess) return nullptr;
    if (hipEventRecord(start, stream) != hipSuccess) {
        (void)hipEventDestroy(start);
        return nullptr;
    }
    return start;
}

void timing_record(const char* name, hipStream_t stream, hipEvent_t start) {
    hipEvent_t end = nullptr;
    if (hipEventCreate(&end) != hipSuccess || hipEventRecord(end, stream) != hipSuccess) {
        if (end != nullptr) (void)hipEventDestroy(end);
        (void)hipEventDestroy(start);
        return;
    }
    std::lock_guard<std::mutex> guard(g_marks_lock);
    g_marks.push_back(Mark{name, start, end});
}

}  // namespace lh

extern "C" {

int lh_timing_enable(int on) {
    lh::g_timing.store(on != 0);
    return LH_OK;
}

int lh_timing_read(const char* kernel, double* total_ms, long long* launches) {
    LH_REQUIRE(kernel != nullptr && total_ms != nullptr && launches != nullptr,
               "lh_timing_read: null argument");
    std::lock_guard<std::mutex> guard(lh::g_marks_lock);
    double total = 0.0;
    long long count = 0;
    for (const lh::Mark& m : lh::g_marks) {
        if (m.name != kernel) continue;
        LH_HIP_CALL(hipEventSynchronize(m.end));
        float ms = 0.0f;
        LH_HIP_CALL(hipEventElapsedTime(&ms, m.start, m.end));
        total += ms;
        ++count;
    }
    *total_ms = total;
    *launches = count;
    return LH_OK;
}

int lh_timing_reset(void) {
    std::lock_guard<std::mutex> guard(lh::g_marks_lock);
    for (lh::Mark& m : lh::g_marks) {
        (void)hipEventSynchronize(m.end);
        (void)hipEventDestroy(m.start);
        (void)hipEventDestroy(m.end);
    }
    lh::g_marks.clear();
    return LH_OK;
}


int lh_graph_begin(void* stream) {
    LH_REQUIRE(stream != nullptr, "lh_graph_begin: capture needs a stream other than the null stream");
    LH_HIP_CALL(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                      hipStreamCaptureModeThreadLocal));
    return LH_OK;
}

int lh_graph_end(void* stream, void** graph_exec) {
    LH_REQUIRE(stream != nullptr && graph_exec != nullptr, "lh_graph_end: null argument");
    hipGraph_t graph = nullptr;
    LH_HIP_CALL(hipStreamEndCapture(static_cast<hipStream_t>(stream), &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t made = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    LH_HIP_CALL(made);
    *graph_exec = exec;
    return LH_OK;
}

int lh_graph_launch(void* graph_exec, void* stream) {
    LH_REQUIRE(graph_exec != nullptr, "lh_graph_launch: null graph");
    LH_HIP_CALL(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec),
                               static_cast<hipStream_t>(stream)));
    return LH_OK;
}

int lh_graph_destroy(void* graph_exec) {
    if (graph_exec != nullptr) LH_HIP_CALL(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
    return LH_OK;
}

const char* lh_last_error(void) { return lh::g_error; }

void lh_clear_error(void) { lh::g_error[0] = '\0'; }

int lh_abi_version(void) { return LH_ABI_VERSION; }

}
