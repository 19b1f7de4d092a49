// gtap_profile.hip -- gtap_visualize_profile: per-worker CSV output of the last launch.
//
// Reference: runtime/detail/thread/gtap_profile_thread.cuh:29-108 (statistics CSV,
// header "warp_id,total_samples,first_activity_ms,last_activity_ms,duration_ms") and
// :355-358 (entry point). Files go to ./profile/<app>_warp_statistics_working.csv with one
// row per wavefront worker and the scheduler counters of that worker.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>

#include "gtap_launch.hpp"

extern "C" void gtap_visualize_profile(const char* app_name) {
    using namespace gtap;
    host::Runtime& R = host::rt();
    const char* app = (app_name && app_name[0]) ? app_name : "gtap";
    if (mkdir("./profile", 0755) != 0 && errno != EEXIST) {
        fprintf(stderr, "GTaP: cannot create ./profile\n");
        return;
    }
    char path[512];
    snprintf(path, sizeof(path), "./profile/%s_warp_statistics_working.csv", app);
    FILE* f = fopen(path, "w");
    if (!f) {
        fprintf(stderr, "GTaP: cannot write %s\n", path);
        return;
    }
    fprintf(f, "warp_id,tasks_executed,rounds,tasks_spawned,pops,steals,steal_attempts,jobs,pool_refills\n");
    const size_t nw = R.stats_host.size() / ST_COUNT;
    for (size_t w = 0; w < nw; ++w) {
        const uint64_t* p = &R.stats_host[w * ST_COUNT];
        fprintf(f, "%zu,%llu,%llu,%llu,%llu,%llu,%llu,%llu,%llu\n", w, (unsigned long long)p[ST_TASKS],
                (unsigned long long)p[ST_ROUNDS], (unsigned long long)p[ST_SPAWNS], (unsigned long long)p[ST_POPS],
                (unsigned long long)p[ST_STEALS], (unsigned long long)p[ST_STEAL_ATTEMPTS],
                (unsigned long long)p[ST_JOBS], (unsigned long long)p[ST_REFILLS]);
    }
    fclose(f);
    printf("Warp statistics saved to: %s\n", path);
}
