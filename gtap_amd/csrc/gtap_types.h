/*
 * gtap_types.h -- plain-C types shared by the host library, the device runtime and
 * the public C ABI (include/gtap.h re-declares the public subset).
 */
#ifndef GTAP_TYPES_H
#define GTAP_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime error codes. 0-8 keep the numbering of the reference
 * (runtime/detail/common/gtap_runtime_error.cuh:7-18); 9+ are additions. */
enum gtap_error_code {
    GTAP_ERROR_NONE = 0,
    GTAP_ERROR_INVALID_QUEUE_IDX = 1,
    GTAP_ERROR_INVALID_QUEUE_IDX_AFTER_JOIN = 2,
    GTAP_ERROR_QUEUE_OVERFLOW = 3,
    GTAP_ERROR_RESULT_HANDLE_OVERFLOW = 4,
    GTAP_ERROR_TASK_ID_POOL_SLOT_BUSY = 5,
    GTAP_ERROR_TASK_ID_POOL_LOW_HEADROOM = 6,
    GTAP_ERROR_GENERATED_TASK_ID_BUFFER_OVERFLOW = 7,
    GTAP_ERROR_INVALID_TASKWAIT = 8,
    GTAP_ERROR_WATCHDOG = 9,
    GTAP_ERROR_INVALID_FUNCTION = 10,
    GTAP_ERROR_INVALID_ARGUMENT = 11,
    GTAP_ERROR_NOT_INITIALIZED = 12,
    GTAP_ERROR_HIP = 13,
    GTAP_ERROR_COUNT
};

/* Device error record, same nine fields as the reference's GTapRuntimeErrorReport
 * (gtap_runtime_error.cuh:20-30). `block_idx`/`thread_idx` name the reporting lane. */
typedef struct gtap_error_report {
    int32_t valid;
    int32_t code;
    int32_t src_line;
    int32_t block_idx;
    int32_t thread_idx;
    int32_t tid;
    int32_t queue_idx;
    int32_t value;
    int32_t limit;
} gtap_error_report;

/* One job of a batched or served stream, 128 bytes. The host writes func/queue/args;
 * the device writes status, result and the timestamps (device wall clock, 100 MHz). */
typedef struct gtap_job {
    uint32_t func;        /* root task function id */
    uint32_t queue;       /* EPAQ queue of the root task */
    uint32_t status;      /* GTAP_JOB_* */
    uint32_t kind;        /* free for the caller (job type tag) */
    int64_t result;       /* value returned by the root task */
    uint64_t t_submit;    /* host: submission time (host clock, ns) */
    uint64_t t_start;     /* device: admission time (wall clock ticks) */
    uint64_t t_end;       /* device: completion time (wall clock ticks) */
    uint64_t t_done_host; /* host: time the completion was seen (host clock, ns) */
    uint64_t reserve_ids; /* served: task ids the job reserves while it runs (0 = none) */
    uint8_t args[64];     /* root frame arguments */
} gtap_job;

enum gtap_job_status { GTAP_JOB_EMPTY = 0, GTAP_JOB_READY = 1, GTAP_JOB_RUNNING = 2, GTAP_JOB_DONE = 3 };

/* Termination modes of one persistent launch. */
enum gtap_term_mode {
    GTAP_TERM_JOBS = 0,     /* exit when every job of the launch has completed
                               (the first-task termination of the reference, per job) */
    GTAP_TERM_QUIESCE = 1,  /* exit when every worker is idle and no job is pending
                               (the all-tasks termination of the reference) */
    GTAP_TERM_SERVE = 2     /* exit when the host has set stop and all admitted jobs completed */
};

/* Per-launch statistics (sums over workers). */
typedef struct gtap_run_stats {
    double kernel_ms;         /* hipEvent time of the persistent launch */
    uint64_t tasks_executed;  /* task invocations, continuations included */
    uint64_t tasks_spawned;   /* spawn calls */
    uint64_t rounds;          /* worker rounds that executed at least one task */
    uint64_t pops;            /* successful pops from a worker's own deque */
    uint64_t steals;          /* successful steals */
    uint64_t steal_attempts;  /* CAS attempts on victims */
    uint64_t jobs_taken;      /* roots admitted from the job table */
    uint64_t pool_refills;    /* id chunks taken from the global pool */
    uint64_t published;       /* task records written out when published (thread mode) */
    uint64_t overflowed;      /* task records written out when pushed out of the LDS window */
    uint64_t suspended;       /* taskwaits that suspended a task */
    uint32_t workers;         /* wavefront workers in the launch */
    uint32_t blocks;          /* workgroups in the launch */
    int32_t error_code;
    int32_t admission_holds;  /* served: admission attempts the id reservation held back
                                 (counted by one wave per pending job) */
} gtap_run_stats;

/* Schedulers of gtap_config.scheduler. */
enum {
    GTAP_SCHED_DEFAULT = 0,
    GTAP_SCHED_WORK_STEALING = 1,
    GTAP_SCHED_GLOBAL_QUEUE = 2,
    GTAP_SCHED_CHASE_LEV = 3
};

/* Tunables of the scheduler (worker count, deque capacity, steal batch width, ...).
 * Zero fields take the defaults. */
typedef struct gtap_config {
    uint64_t pool_bytes;      /* bytes of task records (default and maximum 4 GiB) */
    uint32_t record_stride;   /* bytes per task record (default: max over built-in programs) */
    uint32_t deque_capacity;  /* slots per worker per queue, power of two (default 32768) */
    uint32_t num_queues;      /* EPAQ queues provisioned (default 4) */
    uint32_t blocks;          /* workgroups per launch (default: all co-resident) */
    uint32_t steal_batch;     /* max tasks per steal, <= 64 (default 64) */
    uint32_t probe_lanes;     /* victims probed per steal attempt, <= 64 (default 8) */
    uint32_t max_jobs;        /* job table / anchor capacity (default 65536) */
    uint32_t timeout_ms;      /* device watchdog (default 20000) */
    uint32_t block_workers;   /* workgroups of the block-mode scheduler (default: co-resident) */
    uint32_t idle_sleep;      /* longest backoff of an idle wave between polls, in units of 64
                                 cycles: 1, 2, 4, 8, 16 or 32 (default 32) */
    uint32_t scheduler;       /* thread mode: 0 or 1 = batched work stealing (default), 2 = one
                                 global queue per EPAQ queue, 3 = single-item Chase-Lev pops and
                                 steals (programs built with the alternatives: fib, n-queens) */
    uint32_t epaq_publish;    /* thread mode, several EPAQ queues: 0 = the queues other than the
                                 one the wave's next batch comes from keep their newest <= 64
                                 private entries for the wave's later rounds (default); 1 = they
                                 publish everything at the end of the round */
    uint32_t xcd_steal;       /* thread mode: 1 = half of the steal probes go to victims on the
                                 thief's own XCD (workgroup b runs on XCD b % 8); 0 = all
                                 victims uniform at random (default) */
    uint32_t reserved[2];
} gtap_config;

/* Called by gtap_initialize and gtap_reset with the runtime descriptor (`bytes` = its size): a
 * lowered module copies it to its device symbol, so its host-launched entry kernels reach the
 * runtime's pools. A module loaded after gtap_initialize gets it at the next gtap_reset. */
typedef int (*gtap_entry_setup_fn)(const void* desc, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
