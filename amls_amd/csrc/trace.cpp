// Chrome-trace output of a context: host phase spans and GPU launch spans on
// one timeline.
//
// Follows the reference's task-level tracer (src/core/trace.hpp:82-181):
// enabled by the AMLS_TRACE environment variable (the output path) when the
// context is created, or by amls_hip_trace; events are complete events
// (ph "X") with category, name, lane and start / duration in microseconds,
// written as {"traceEvents": [...]} for chrome://tracing or Perfetto.  The
// host side records one lane (tid 0: set-up, elimination, condensation,
// reduced solve, back-transformation, subspace iteration); the GPU side adds
// every launch of the engine, measured by HIP events and placed on the host
// clock through one base event (tid 1: the main stream, tid 2: the second
// stream), each with the solve phase it belongs to.
#include <cstdio>

#include "engine.hpp"

namespace amls {

namespace {

int64_t usOf(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration_cast<std::chrono::microseconds>(t.time_since_epoch()).count();
}

void jsonString(std::FILE* f, const std::string& s) {
    std::fputc('"', f);
    for (char c : s) {
        if (c == '"' || c == '\\') std::fputc('\\', f);
        if (static_cast<unsigned char>(c) >= 0x20) std::fputc(c, f);
    }
    std::fputc('"', f);
}

}  // namespace

void Ctx::traceHost(const char* cat, const char* name, std::chrono::steady_clock::time_point a,
                    std::chrono::steady_clock::time_point b) {
    if (!kstats.trace) return;
    kstats.spans.push_back({cat, name, "", 0, usOf(a), usOf(b)});
}

// writes every span recorded so far (launches still in flight are resolved
// first) and returns their number, or -1 when the file cannot be written
int64_t Ctx::traceWrite(const std::string& path) {
    if (gpu) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamSynchronize(stream2);
        if (stream3) (void)hipStreamSynchronize(stream3);
        kstats.resolve(true);
    }
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return -1;
    std::fprintf(f, "{\"traceEvents\":[\n");
    const char* lanes[3] = {"host", "gpu stream 1", "gpu stream 2"};
    for (int t = 0; t < 3; ++t)
        std::fprintf(f,
                     "{\"ph\":\"M\",\"name\":\"thread_name\",\"pid\":0,\"tid\":%d,"
                     "\"args\":{\"name\":\"%s\"}},\n",
                     t, lanes[t]);
    int64_t n = 0;
    for (const auto& e : kstats.spans) {
        if (n) std::fprintf(f, ",\n");
        std::fprintf(f, "{\"ph\":\"X\",\"cat\":");
        jsonString(f, e.cat);
        std::fprintf(f, ",\"name\":");
        jsonString(f, e.name);
        std::fprintf(f, ",\"pid\":0,\"tid\":%d,\"ts\":%lld,\"dur\":%lld", e.tid,
                     static_cast<long long>(e.t0), static_cast<long long>(e.t1 - e.t0));
        if (!e.phase.empty()) {
            std::fprintf(f, ",\"args\":{\"phase\":");
            jsonString(f, e.phase);
            std::fprintf(f, "}");
        }
        std::fprintf(f, "}");
        ++n;
    }
    std::fprintf(f, "\n]}\n");
    std::fclose(f);
    return n;
}

}  // namespace amls
