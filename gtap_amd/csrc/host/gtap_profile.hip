// gtap_profile.hip -- gtap_visualize_profile: CSV output of the last launch.
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

}  // namespace

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
        return;
    }
    const uint32_t nw = R.last_nworkers, M = R.prof_max;
    std::vector<uint64_t> hav((size_t)nw * M), wrk((size_t)nw * M);
    std::vector<int32_t> cnt((size_t)nw * M);
    if (hipMemcpy(hav.data(), R.prof_having, hav.size() * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(wrk.data(), R.prof_working, wrk.size() * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(cnt.data(), R.prof_counts, cnt.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        fprintf(stderr, "GTaP: cannot read the profile buffers\n");
        return;
    }
    uint64_t min_t = UINT64_MAX, max_t = 0;
    for (size_t i = 0; i < hav.size(); ++i) {
        if (hav[i]) { min_t = std::min(min_t, hav[i]); max_t = std::max(max_t, hav[i]); }
        if (wrk[i]) { min_t = std::min(min_t, wrk[i]); max_t = std::max(max_t, wrk[i]); }
    }
    if (min_t == UINT64_MAX) {
        printf("GTaP: no profile samples recorded\n");
        return;
    }
    auto rel = [min_t](uint64_t t) { return (double)(t - min_t) / 1e6; };
    // having-task timeline and statistics
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
    // working timeline and statistics
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
    printf("GTaP profile: %u workers, %.3f ms span, CSV files in ./profile/%s_warp_*.csv\n", nw,
           (double)(max_t - min_t) / 1e6, app);
}
