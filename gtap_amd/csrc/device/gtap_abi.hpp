// gtap_abi.hpp -- the compiler <-> runtime device ABI used by lowered task functions.
//
// A `#pragma gtap function` becomes a switch-based state machine over the resume state
// (the contract of the reference's clang-gtap; runtime side in
// runtime/detail/thread/gtap_runtime_thread.cuh:663-847). The names below keep the
// reference's `__gtap_*` spelling; their arguments follow this runtime's design:
// tasks are named by a function id (a `switch` dispatch instead of an indirect call,
// which would block inlining on AMDGPU), and a spawn names the byte offset of the
// parent's result slot instead of retaining the child frame.
//
//   reference                                   here
//   __gtap_get_task_state(tid)                  __gtap_get_task_state(L)
//   __gtap_spawn_task(ctx, tid, &cnt, f, q,     __gtap_spawn_task(L, f, q, res_off, args)
//                     &out_tid, retain)         __gtap_spawn_task_detached(L, f, q, args)
//   __gtap_set_state_for_join(tid, cnt, s, q)   __gtap_set_state_for_join(L, s, q)
//   __gtap_finish_task(tid, ctx)                __gtap_finish_task(L) / __gtap_return(L, v)
//   __gtap_get_task_data(tid)                   __gtap_get_task_data(L)
//
// The additive join has no reference counterpart: a parent that only sums its children's
// integer results spawns them with __gtap_spawn_task_sum and waits with
// __gtap_set_state_for_join_sum; after resuming, __gtap_get_join_sum(L) is the sum. Each
// child's result then travels in the atomic that counts the join down (thread-mode lane
// tasks; bounds in gtap_dev.hpp).
#pragma once

#include "gtap_block.hpp"
#include "gtap_thread.hpp"

template <class L>
__device__ __forceinline__ uint32_t __gtap_get_task_state(const L& lane) {
    return lane.get_state();
}
template <class L, class A>
__device__ __forceinline__ uint32_t __gtap_spawn_task(L& lane, uint32_t func, uint32_t queue, int res_off,
                                                      const A& args) {
    return lane.spawn(func, queue, res_off, args);
}
template <class L, class A>
__device__ __forceinline__ uint32_t __gtap_spawn_task_detached(L& lane, uint32_t func, uint32_t queue, const A& args) {
    return lane.spawn_detached(func, queue, args);
}
template <class L>
__device__ __forceinline__ bool __gtap_set_state_for_join(L& lane, uint32_t next_state, uint32_t queue_after) {
    return lane.taskwait(next_state, queue_after);
}
template <class L, class A>
__device__ __forceinline__ uint32_t __gtap_spawn_task_sum(L& lane, uint32_t func, uint32_t queue, const A& args) {
    return lane.spawn_sum(func, queue, args);
}
template <class L>
__device__ __forceinline__ bool __gtap_set_state_for_join_sum(L& lane, uint32_t next_state, uint32_t queue_after) {
    return lane.taskwait_sum(next_state, queue_after);
}
template <class L>
__device__ __forceinline__ int64_t __gtap_get_join_sum(const L& lane) {
    return lane.join_sum();
}
template <class L>
__device__ __forceinline__ void __gtap_finish_task(L& lane) {
    lane.finish();
}
template <class L, class T>
__device__ __forceinline__ void __gtap_return(L& lane, T value) {
    lane.ret(value);
}
template <class L>
__device__ __forceinline__ uint8_t* __gtap_get_task_data(L& lane) {
    return lane.rec + ::gtap::kHdrBytes;
}

// Block mode (gtap_runtime_block.cuh:441-576): the same calls on a workgroup task. A spawn may
// come from any subset of the threads; the queue is ignored (one deque per workgroup, as in
// the reference's block mode); taskwait, return and finish are workgroup-collective.
template <class P, class A>
__device__ __forceinline__ uint32_t __gtap_spawn_task(::gtap::BlockTask<P>& blk, uint32_t func, uint32_t /*queue*/,
                                                      int res_off, const A& args) {
    return blk.spawn(func, res_off, args);
}
template <class P, class A>
__device__ __forceinline__ uint32_t __gtap_spawn_task_detached(::gtap::BlockTask<P>& blk, uint32_t func,
                                                               uint32_t /*queue*/, const A& args) {
    return blk.spawn_detached(func, args);
}
template <class P>
__device__ __forceinline__ bool __gtap_set_state_for_join(::gtap::BlockTask<P>& blk, uint32_t next_state,
                                                          uint32_t /*queue_after*/) {
    return blk.taskwait(next_state);
}
// (reference: __gtap_set_state_for_join_block(tid, ctx, next_state, unused), the child count
// taken from the workgroup's context)
template <class P>
__device__ __forceinline__ bool __gtap_set_state_for_join_block(::gtap::BlockTask<P>& blk, uint32_t next_state) {
    return blk.taskwait(next_state);
}
