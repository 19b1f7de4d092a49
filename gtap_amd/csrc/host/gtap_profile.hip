// gtap_profile.hip -- gtap_visualize_profile, gtap_visualize_having_task_time,
// gtap_visualize_working_time: console timelines and CSV output of the last launch.
//
// With timeline profiling on (gtap_set_profiling), the four files of the reference's profiler
// (runtime/detail/thread/gtap_profile_thread.cuh:29-108 and 188-268), same names and columns:
//   ./profile/<app>_warp_timeline_having_task.csv
//       warp_id,time_index,timestamp_ns,relative_time_ms,state,state_description
//   ./profile/<app>_warp_statistics_having_task.csv
//       warp_id,total_samples,first_activity_ms,last_activity_ms,duration_ms
//   ./profile/<app>_warp_timeline_working.csv
//       warp_id,time_index,timestamp_ns,relative_time_ms,state,state_description,tasks_in_batch
//   ./profile/<app>_warp_statistics_working.csv
//       warp_id,total_samples,first_activity_ms,last_activity_ms,duration_ms,total_tasks
// A "warp" row is one wavefront worker (thread mode) or one workgroup (block mode).
// Always written: ./profile/<app>_warp_counters.csv with the scheduler counters per worker.
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <vector>

#include "gtap_launch.hpp"

namespace {

bool ensure_dir() {
    if (mkdir("./profile", 0755) != 0 && errno != EEXIST) {
        fprintf(stderr, "GTaP: cannot create ./profile\n");
        return false;
    }
    return true;
}

FILE* open_csv(const char* app, const char* suffix) {
    char path[512];
    snprintf(path, sizeof(path), "./profile/%s_%s.csv", app, suffix);
    FILE* f = fopen(path, "w");
    if (!f) fprintf(stderr, "GTaP: cannot write %s\n", path);
    return f;
}

// The last launch's timeline samples, read back once per call. Each sample list of worker w
// is hav/wrk[w * M + i] for i < M up to the first zero; even entries open an interval (has a
// task / runs a batch), odd entries close it.
struct Samples {
    uint32_t nw = 0, M = 0;
    std::vector<uint64_t> hav, wrk;
    std::vector<int32_t> cnt;
};

// 0: read; 1: profiling is off or there was no launch; 2: read error.
int read_samples(Samples& S) {
    gtap::host::Runtime& R = gtap::host::rt();
    if (!R.prof_max || !R.last_nworkers) return 1;
    S.nw = R.last_nworkers;
    S.M = R.prof_max;
    S.hav.resize((size_t)S.nw * S.M);
    S.wrk.resize((size_t)S.nw * S.M);
    S.cnt.resize((size_t)S.nw * S.M);
    if (hipMemcpy(S.hav.data(), R.prof_having, S.hav.size() * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(S.wrk.data(), R.prof_working, S.wrk.size() * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(S.cnt.data(), R.prof_counts, S.cnt.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        fprintf(stderr, "GTaP: cannot read the profile buffers\n");
        return 2;
    }
    return 0;
}

// Time range of one sample array; false when it holds no sample.
bool sample_range(const Samples& S, const std::vector<uint64_t>& v, uint64_t& lo, uint64_t& hi, uint64_t& count) {
    lo = UINT64_MAX;
    hi = 0;
    count = 0;
    for (uint32_t w = 0; w < S.nw; ++w)
        for (uint32_t i = 0; i < S.M && v[(size_t)w * S.M + i]; ++i) {
            const uint64_t t = v[(size_t)w * S.M + i];
            lo = std::min(lo, t);
            hi = std::max(hi, t);
            ++count;
        }
    return count != 0;
}

// Text timeline of the first (up to) ten workers over 80 columns: each sample marks its
// column with `open` (even samples) or 'N' (odd samples); the reference's console view
// (runtime/detail/thread/gtap_profile_thread.cuh:157-180, 323-346).
void print_timeline(const char* title, const Samples& S, const std::vector<uint64_t>& v, uint64_t lo, uint64_t hi,
                    uint64_t count, char open) {
    constexpr int kWidth = 80;
    printf("\n=== Warp %s Time Visualization ===\n", title);
    printf("Total samples: %llu\n", (unsigned long long)count);
    printf("Time range: %.3f ms to %.3f ms\n", (double)lo / 1e6, (double)hi / 1e6);
    printf("Duration: %.3f ms\n", (double)(hi - lo) / 1e6);
    const uint32_t shown = std::min<uint32_t>(S.nw, 10);
    printf("\nTimeline visualization (first %u warps):\n", shown);
    char ruler[kWidth + 1];
    for (int c = 0; c < kWidth; ++c) ruler[c] = (c % 10 == 0) ? '|' : '-';
    ruler[kWidth] = '\0';
    printf("Warp  |%s\n", ruler);
    const uint64_t span = hi > lo ? hi - lo : 1;
    for (uint32_t w = 0; w < shown; ++w) {
        char row[kWidth + 1];
        for (int c = 0; c < kWidth; ++c) row[c] = ' ';
        row[kWidth] = '\0';
        for (uint32_t i = 0; i < S.M && v[(size_t)w * S.M + i]; ++i) {
            const uint64_t col = (v[(size_t)w * S.M + i] - lo) * (uint64_t)kWidth / span;
            if (col < (uint64_t)kWidth) row[col] = (i % 2 == 0) ? open : 'N';
        }
        printf("%5u |%s\n", w, row);
    }
}

// Reads the samples and creates ./profile; false (after a message) when there is nothing to show.
bool prologue(const char* app_name, const char*& app, Samples& S, const char* what) {
    app = (app_name && app_name[0]) ? app_name : "gtap";
    const int st = read_samples(S);
    if (st == 1) {
        printf("GTaP: timeline profiling is off or nothing ran (gtap_set_profiling); no %s timeline\n", what);
        fflush(stdout);
        return false;
    }
    return st == 0 && ensure_dir();
}

}  // namespace

// The having-task view: printed timeline ('H' = got a task, 'N' = ran out) and
// ./profile/<app>_warp_{timeline,statistics}_having_task.csv, times relative to the first
// having-task sample (runtime/detail/thread/gtap_profile_thread.cuh:29-108, 110-186).
extern "C" void gtap_visualize_having_task_time(const char* app_name) {
    const char* app = nullptr;
    Samples S;
    if (!prologue(app_name, app, S, "having-task")) return;
    uint64_t lo, hi, count;
    if (!sample_range(S, S.hav, lo, hi, count)) {
        printf("No having-task time data recorded.\n");
        fflush(stdout);
        return;
    }
    print_timeline("Having-Task", S, S.hav, lo, hi, count, 'H');
    auto rel = [lo](uint64_t t) { return (double)(t - lo) / 1e6; };
    const uint32_t nw = S.nw, M = S.M;
    const std::vector<uint64_t>& hav = S.hav;
    if (FILE* f = open_csv(app, "warp_timeline_having_task")) {
        fprintf(f, "warp_id,time_index,timestamp_ns,relative_time_ms,state,state_description\n");
        for (uint32_t w = 0; w < nw; ++w)
            for (uint32_t i = 0; i < M && hav[(size_t)w * M + i]; ++i) {
                const uint64_t t = hav[(size_t)w * M + i];
                fprintf(f, "%u,%u,%llu,%.6f,%d,%s\n", w, i, (unsigned long long)t, rel(t), (i % 2 == 0) ? 1 : 0,
                        (i % 2 == 0) ? "Having" : "NotHaving");
            }
        fclose(f);
    }
    if (FILE* f = open_csv(app, "warp_statistics_having_task")) {
        fprintf(f, "warp_id,total_samples,first_activity_ms,last_activity_ms,duration_ms\n");
        for (uint32_t w = 0; w < nw; ++w) {
            uint32_t n = 0, having = 0;
            uint64_t first = UINT64_MAX, last = 0;
            for (uint32_t i = 0; i < M && hav[(size_t)w * M + i]; ++i, ++n) {
                const uint64_t t = hav[(size_t)w * M + i];
                first = std::min(first, t);
                last = std::max(last, t);
                if (i % 2 == 0) ++having;
            }
            if (!n) fprintf(f, "%u,0,0.0,0.0,0.0\n", w);
            else fprintf(f, "%u,%u,%.6f,%.6f,%.6f\n", w, having, rel(first), rel(last), (double)(last - first) / 1e6);
        }
        fclose(f);
    }
    printf("Warp timelines saved to: ./profile/%s_warp_{timeline,statistics}_having_task.csv\n", app);
    fflush(stdout);
}

// The working view: printed timeline ('W' = a batch starts, 'N' = it ends) and
// ./profile/<app>_warp_{timeline,statistics}_working.csv with the batch sizes, times relative
// to the first working sample (runtime/detail/thread/gtap_profile_thread.cuh:188-268, 270-352).
extern "C" void gtap_visualize_working_time(const char* app_name) {
    const char* app = nullptr;
    Samples S;
    if (!prologue(app_name, app, S, "working")) return;
    uint64_t lo, hi, count;
    if (!sample_range(S, S.wrk, lo, hi, count)) {
        printf("No working time data recorded.\n");
        fflush(stdout);
        return;
    }
    print_timeline("Working", S, S.wrk, lo, hi, count, 'W');
    auto rel = [lo](uint64_t t) { return (double)(t - lo) / 1e6; };
    const uint32_t nw = S.nw, M = S.M;
    const std::vector<uint64_t>& wrk = S.wrk;
    const std::vector<int32_t>& cnt = S.cnt;
    if (FILE* f = open_csv(app, "warp_timeline_working")) {
        fprintf(f, "warp_id,time_index,timestamp_ns,relative_time_ms,state,state_description,tasks_in_batch\n");
        for (uint32_t w = 0; w < nw; ++w)
            for (uint32_t i = 0; i < M && wrk[(size_t)w * M + i]; ++i) {
                const uint64_t t = wrk[(size_t)w * M + i];
                fprintf(f, "%u,%u,%llu,%.6f,%d,%s,%d\n", w, i, (unsigned long long)t, rel(t), (i % 2 == 0) ? 1 : 0,
                        (i % 2 == 0) ? "Working" : "NotWorking", cnt[(size_t)w * M + i]);
            }
        fclose(f);
    }
    if (FILE* f = open_csv(app, "warp_statistics_working")) {
        fprintf(f, "warp_id,total_samples,first_activity_ms,last_activity_ms,duration_ms,total_tasks\n");
        for (uint32_t w = 0; w < nw; ++w) {
            uint32_t n = 0, pre = 0;
            long long total_tasks = 0;
            uint64_t first = UINT64_MAX, last = 0;
            for (uint32_t i = 0; i < M && wrk[(size_t)w * M + i]; ++i, ++n) {
                const uint64_t t = wrk[(size_t)w * M + i];
                first = std::min(first, t);
                last = std::max(last, t);
                if (i % 2 == 0) ++pre;
                else total_tasks += cnt[(size_t)w * M + i];
            }
            if (!n) fprintf(f, "%u,0,0.0,0.0,0.0,0\n", w);
            else
                fprintf(f, "%u,%u,%.6f,%.6f,%.6f,%lld\n", w, pre, rel(first), rel(last), (double)(last - first) / 1e6,
                        total_tasks);
        }
        fclose(f);
    }
    printf("Warp timelines saved to: ./profile/%s_warp_{timeline,statistics}_working.csv\n", app);
    fflush(stdout);
}

// Both views, plus ./profile/<app>_warp_counters.csv with the scheduler counters per worker
// (written also when timeline profiling is off).
extern "C" void gtap_visualize_profile(const char* app_name) {
    using namespace gtap;
    host::Runtime& R = host::rt();
    const char* app = (app_name && app_name[0]) ? app_name : "gtap";
    if (!ensure_dir()) return;
    if (R.initialized && host::fetch_worker_stats() != 0) return;

    if (FILE* f = open_csv(app, "warp_counters")) {
        fprintf(f, "warp_id,tasks_executed,rounds,tasks_spawned,pops,steals,steal_attempts,jobs,pool_refills\n");
        const size_t nw = R.stats_host.size() / ST_COUNT;
        for (size_t w = 0; w < nw; ++w) {
            const uint64_t* p = &R.stats_host[w * ST_COUNT];
            fprintf(f, "%zu,%llu,%llu,%llu,%llu,%llu,%llu,%llu,%llu\n", w, (unsigned long long)p[ST_TASKS],
                    (unsigned long long)p[ST_ROUNDS], (unsigned long long)p[ST_SPAWNS],
                    (unsigned long long)p[ST_POPS], (unsigned long long)p[ST_STEALS],
                    (unsigned long long)p[ST_STEAL_ATTEMPTS], (unsigned long long)p[ST_JOBS],
                    (unsigned long long)p[ST_REFILLS]);
        }
        fclose(f);
    }
    if (!R.prof_max || !R.last_nworkers) {
        printf("GTaP: timeline profiling is off (gtap_set_profiling); wrote %s_warp_counters.csv\n", app);
        fflush(stdout);
        return;
    }
    gtap_visualize_having_task_time(app);
    gtap_visualize_working_time(app);
}
