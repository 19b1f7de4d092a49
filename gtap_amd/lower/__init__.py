"""gtap_amd.lower -- source-to-source lowering of `#pragma gtap` task programs to HIP for this
runtime (the role of the reference's clang-gtap, which is not part of the reference tree;
its contract is the device ABI of runtime/detail/thread/gtap_runtime_thread.cuh:663-847 and
runtime/detail/block/gtap_runtime_block.cuh:441-602).

Input: a CUDA/HIP source in the reference's dialect, unmodified --

    #include "gtap_thread.cuh"            // or gtap_block.cuh: block mode
    __device__ int d_result;
    __device__ __constant__ int d_cutoff; // set by the host with hipMemcpyToSymbol

    #pragma gtap function
    __device__ int fib(int n) {
        if (n < d_cutoff) return n;
        int a, b;
        #pragma gtap task queue(n < 10 ? 1 : 0)
        a = fib(n - 1);
        #pragma gtap task
        b = fib(n - 2);
        #pragma gtap taskwait queue(2)
        return a + b;
    }
    __global__ void exec_kernel(int n) {
        #pragma gtap entry
        d_result = fib(n);
    }

plus the compile-time switches the reference takes on its command line (`-D
GTAP_NUM_QUEUES=3`, `-D GTAP_TERMINATE_ON_FIRST_TASK_FINISH`, program macros), and
`GTAP_WAVES_PER_SIMD=N`, which asks for at least N waves per SIMD in the scheduler kernels (the
compiler keeps their registers within that occupancy, spilling if it must).

Output: one .hip file that keeps the source's device side and its host-launch convention:
  * the source's preprocessor definitions, types, file-scope `__device__` / `__constant__`
    variables and device helper functions, at file scope under their own names (the host
    sets and reads the variables with hipMemcpyToSymbol / hipMemcpyFromSymbol);
  * per task function f: `f__args` / `f__frame` (parameters, then every local) and
    `f__task`, a switch over the resume state in which every local lives in the frame struct
    `fr` (protothread style). A spawn names the child's function id and the byte offset of
    its result slot in the parent's frame. A spawn that a later taskwait of the function
    joins is a joined child; one that no taskwait follows is detached (the reference's
    semantics for children of a task that returns without a taskwait). A program with no
    detached spawn ends on the job count (its root ends last), one with detached spawns on
    quiescence, unless GTAP_TERMINATE_ON_FIRST_TASK_FINISH asks for the root's end. `f__seq`
    is an ordinary device function for direct calls;
  * every `__global__` kernel with `#pragma gtap entry`, under its own name and parameters,
    launchable by the host as `kernel<<<GTAP_GRID_SIZE, GTAP_BLOCK_SIZE>>>(...)`: the
    kernel's own statements run as written, the entry call runs the persistent scheduler
    on the launch's wavefronts (thread mode) or workgroups (block mode), and the entry's
    lhs (`d_result`) receives the root's result when the launch's work is complete. One
    launch may follow each gtap_initialize / gtap_reset. GTAP_GRID_SIZE is the co-resident
    grid (`gtap_lowered_<name>_launch_dims`), GTAP_BLOCK_SIZE 256;
  * `extern "C" int gtap_entry_<kernel>(<params>, <result>* result, gtap_run_stats*)`, the
    same root as one persistent launch by the library (the kernel's other statements are
    not run), and `extern "C" int gtap_launch_<kernel>(grid, block, <params>)`, the host
    launch above as a C function.

Memory that tasks share: this runtime publishes work without cache fences (one L2 per XCD),
so every access through a pointer in task functions, their device helpers and entry kernels
is rewritten to a coherent (agent-scope) load or store: `a[i]` read becomes
`gtap::cld(&(a[i]))`, `a[i] = v` becomes `gtap::coh(a[i]) = v`. Accesses that cannot be made
coherent this way (inline asm, memcpy/memset on shared pointers) are rejected.

Restrictions (reported as LoweringError): parameters are scalars or pointers of 4 or 8
bytes; locals live across a taskwait are 4- or 8-byte scalars, pointers, or arrays of them
without initialisers; a task's `lhs` is a local or an element of a local array of the
callee's return size, and needs a later taskwait. In block mode the locals that live across
a taskwait must hold the same value on every thread (thread 0's values are saved).
"""
from __future__ import annotations

import os
import re
import tempfile
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

from . import _clang as cl

PRAGMA_RE = re.compile(r"^[ \t]*#[ \t]*pragma[ \t]+gtap[ \t]+(function|task|taskwait|entry)\b([^\n]*)$", re.M)
QUEUE_RE = re.compile(r"queue\s*\((.*)\)\s*$")
INCLUDE_RE = re.compile(r"^[ \t]*#[ \t]*include[^\n]*$", re.M)
DIRECTIVE_RE = re.compile(r"^[ \t]*#(?:[^\n]*\\\n)*[^\n]*$", re.M)
DEFINE_RE = re.compile(r"^[ \t]*#[ \t]*define[ \t]+([A-Za-z_]\w*)(?:[ \t]+([^\n]*))?$", re.M)
# includes that name the CUDA runtime or the reference's runtime: replaced by this runtime's headers
DROP_INCLUDE_RE = re.compile(r"(cuda[\w./]*\.h|gtap_\w+\.cuh|cooperative_groups|device_launch_parameters|curand)")

PRELUDE = """#define __device__ __attribute__((annotate("gtap_device")))
#define __global__ __attribute__((annotate("gtap_global")))
#define __constant__ __attribute__((annotate("gtap_constant")))
#define __managed__
#define __host__
#define __shared__
#define __forceinline__ inline
#define __noinline__
#define __restrict__
#define __launch_bounds__(...)
typedef signed char int8_t; typedef unsigned char uint8_t; typedef short int16_t;
typedef unsigned short uint16_t; typedef int int32_t; typedef unsigned int uint32_t;
typedef long int64_t; typedef unsigned long uint64_t; typedef unsigned long size_t;
typedef long ptrdiff_t; typedef long intptr_t; typedef unsigned long uintptr_t;
#define NULL 0
#ifndef GTAP_BLOCK_SIZE
#define GTAP_BLOCK_SIZE 256
#endif
#ifndef GTAP_GRID_SIZE
#define GTAP_GRID_SIZE 1024
#endif
#ifndef GTAP_WARP_SIZE
#define GTAP_WARP_SIZE 64
#endif
#define GTAP_NUM_WARPS (GTAP_BLOCK_SIZE / GTAP_WARP_SIZE)
#define INT_MAX 2147483647
#define INT_MIN (-INT_MAX - 1)
#define UINT_MAX 4294967295U
#define LLONG_MAX 9223372036854775807LL
#define LLONG_MIN (-LLONG_MAX - 1LL)
#define ULLONG_MAX 18446744073709551615ULL
#define DBL_MAX 1.79769313486231570815e+308
#define FLT_MAX 3.40282346638528859812e+38F
struct __gtap_dim3 { unsigned x, y, z; };
extern const __gtap_dim3 threadIdx, blockIdx, blockDim, gridDim;
extern const int warpSize;
void __syncthreads(); void __threadfence(); void __threadfence_block(); void __threadfence_system();
int printf(const char*, ...);
void* memcpy(void*, const void*, size_t); void* memset(void*, int, size_t); void* memmove(void*, const void*, size_t);
int atomicAdd(int*, int); unsigned atomicAdd(unsigned*, unsigned);
unsigned long long atomicAdd(unsigned long long*, unsigned long long); float atomicAdd(float*, float);
double atomicAdd(double*, double);
int atomicSub(int*, int); unsigned atomicSub(unsigned*, unsigned);
int atomicMin(int*, int); unsigned atomicMin(unsigned*, unsigned);
int atomicMax(int*, int); unsigned atomicMax(unsigned*, unsigned);
int atomicCAS(int*, int, int); unsigned atomicCAS(unsigned*, unsigned, unsigned);
unsigned long long atomicCAS(unsigned long long*, unsigned long long, unsigned long long);
int atomicExch(int*, int); unsigned atomicExch(unsigned*, unsigned);
int atomicOr(int*, int); unsigned atomicOr(unsigned*, unsigned);
int atomicAnd(int*, int); unsigned atomicAnd(unsigned*, unsigned);
double fma(double, double, double); float fmaf(float, float, float);
double sqrt(double); float sqrtf(float); double fabs(double); float fabsf(float);
double exp(double); double log(double); double pow(double, double); double floor(double); double ceil(double);
int abs(int); long long llabs(long long);
int min(int, int); unsigned min(unsigned, unsigned); long long min(long long, long long);
float min(float, float); double min(double, double);
int max(int, int); unsigned max(unsigned, unsigned); long long max(long long, long long);
float max(float, float); double max(double, double);
long long __double_as_longlong(double); double __longlong_as_double(long long);
int __float_as_int(float); float __int_as_float(int);
int __popc(unsigned); int __popcll(unsigned long long); int __clz(int); int __clzll(long long);
int __ffs(int); int __ffsll(long long); unsigned __brev(unsigned);
long long clock64(); long long clock();
template <class T> T gtap_coherent_load(const T* p);
template <class T> void gtap_coherent_store(T* p, T v);
int load_L2(int*); unsigned load_L2(unsigned*); uint16_t load_L2_u16t(uint16_t*); int load_L2_acquire(int*);
void* load_L2_ptr(void**); template <class T> T* load_L2_ptr(T**);
void store_L2(int*, int); void store_L2(unsigned*, unsigned); void store_L2_u16t(uint16_t*, uint16_t);
void store_L2_release(int*, int); void store_L2_ptr(void**, void*); template <class T> void store_L2_ptr(T**, T*);
template <class T> void store_struct_L2(T*, const T&);
int atomicCAS_acquire(int*, int, int); void lock(int*); void unlock(int*);
unsigned get_lane_id(); unsigned get_warp_id_in_block();
"""

# libclang cursor kinds used below (clang-c/Index.h)
STRUCT_DECL, UNION_DECL, CLASS_DECL, ENUM_DECL, TYPEDEF_DECL, FUNCTION_TEMPLATE, TYPE_ALIAS_DECL = 2, 3, 4, 5, 20, 30, 36
UNEXPOSED_EXPR = 100
PAREN_EXPR = 111
UNARY_OPERATOR = 112
COMPOUND_ASSIGN = 115
ASM_STMTS = (215, 216)
ANNOTATE_ATTR = 406
EXPR_KINDS = range(100, 200)
# statements that make what they contain conditional or repeated: label, case, default, if,
# switch, while, do, for, range-for
CONTROL_KINDS = (201, 203, 204, 205, 206, 207, 208, 209, 225)
# 4-byte integer types whose sums the additive join carries exactly
INT32_TYPES = {"int", "signed", "signed int", "unsigned", "unsigned int", "int32_t", "uint32_t"}


class LoweringError(Exception):
    pass


@dataclass
class Pragma:
    kind: str
    start: int      # offset of the line start
    end: int        # offset just past the line
    queue: str      # queue expression text ("0" when absent)


@dataclass
class Var:
    name: str       # field name in the frame
    decl: Tuple     # (kind, offset[, name]) identity of the declaration
    type: str
    size: int
    elem_type: str = ""
    elem_size: int = 0
    array_len: int = -1
    is_param: bool = False

    def decl_text(self) -> str:
        if self.array_len >= 0:
            return f"{self.elem_type} {self.name}[{self.array_len}];"
        return f"{self.type} {self.name};"


@dataclass
class TaskFn:
    name: str
    node: cl.Node
    body: cl.Node
    ret_type: str
    ret_size: int
    params: List[Var] = field(default_factory=list)
    locals: List[Var] = field(default_factory=list)
    result_targets: set = field(default_factory=set)
    func_id: int = 0
    taskwaits: List[int] = field(default_factory=list)  # pragma offsets
    lds_locals: set = field(default_factory=set)        # offsets of __shared__ local declarations
    sum_vars: Optional[set] = None                      # result locals carried by the additive join


def _stmt_end(text: str, pos: int) -> int:
    """Offset just past the ';' that ends the expression statement reaching `pos`."""
    i = pos
    while i < len(text) and text[i] in " \t\r\n":
        i += 1
    if i >= len(text) or text[i] != ";":
        raise LoweringError(f"expected ';' at offset {pos}")
    return i + 1


def _annotations(n: cl.Node) -> set:
    return {c.spelling for c in n.children if c.kind == ANNOTATE_ATTR}


def _is_device(n: cl.Node) -> bool:
    return "gtap_device" in _annotations(n)


def _is_global(n: cl.Node) -> bool:
    return "gtap_global" in _annotations(n)


def _is_constant(n: cl.Node) -> bool:
    return "gtap_constant" in _annotations(n)


def _body(n: cl.Node) -> Optional[cl.Node]:
    for c in n.children:
        if c.kind == cl.COMPOUND_STMT:
            return c
    return None


def _parents(root: cl.Node) -> Dict[int, cl.Node]:
    par: Dict[int, cl.Node] = {}
    for n in root.walk():
        for c in n.children:
            par[id(c)] = n
    return par


def _strip(n: cl.Node) -> cl.Node:
    """The expression under implicit casts and parentheses."""
    while n.kind in (UNEXPOSED_EXPR, PAREN_EXPR) and len(n.children) == 1:
        n = n.children[0]
    return n


def _unop(text: str, n: cl.Node) -> str:
    """The operator of a UnaryOperator cursor (libclang does not spell it): the text before
    its operand (prefix) or after it (postfix)."""
    if not n.children:
        return ""
    c = n.children[0]
    pre = text[n.start:c.start].strip()
    return pre if pre else text[c.end:n.end].strip()


def _is_array_type(t: str) -> bool:
    return "[" in t and not t.rstrip().endswith("*")


class _Lowerer:
    def __init__(self, text: str, path: str, name: str, defines: Optional[Dict[str, str]] = None,
                 mode: Optional[str] = None, coherent: bool = True):
        # one character per byte (latin-1 view of the UTF-8 source): libclang's offsets are byte
        # offsets, so comments in any script keep the AST extents aligned with this text
        self.text = text.encode("utf-8").decode("latin-1")
        self.path = path
        self.name = re.sub(r"\W", "_", name)
        self.coherent = coherent
        self.cli_defines: Dict[str, str] = dict(defines or {})
        self.pragmas: List[Pragma] = []
        text = self.text
        for m in PRAGMA_RE.finditer(text):
            rest = m.group(2).strip()
            q = QUEUE_RE.match(rest)
            self.pragmas.append(Pragma(m.group(1), m.start(), m.end(), q.group(1).strip() if q else "0"))
        src_defs = {m.group(1): (m.group(2) or "").strip() for m in DEFINE_RE.finditer(text)}
        alldefs = {**src_defs, **self.cli_defines}
        try:
            self.num_queues = int(alldefs.get("GTAP_NUM_QUEUES", "1") or "1")
        except ValueError:
            raise LoweringError(f"GTAP_NUM_QUEUES must be an integer ({alldefs['GTAP_NUM_QUEUES']!r})")
        if not 1 <= self.num_queues <= 8:
            raise LoweringError("GTAP_NUM_QUEUES must be in [1, 8]")
        self.first_task_termination = "GTAP_TERMINATE_ON_FIRST_TASK_FINISH" in alldefs
        self.detached_spawns = 0
        # minimum waves per SIMD of the scheduler kernels (the compiler keeps registers within
        # it, spilling if it must); absent: the compiler's choice
        self.waves_per_simd: Optional[int] = None
        if "GTAP_WAVES_PER_SIMD" in alldefs:
            try:
                self.waves_per_simd = int(alldefs["GTAP_WAVES_PER_SIMD"])
            except ValueError:
                raise LoweringError(f"GTAP_WAVES_PER_SIMD must be an integer ({alldefs['GTAP_WAVES_PER_SIMD']!r})")
            if not 1 <= self.waves_per_simd <= 8:
                raise LoweringError("GTAP_WAVES_PER_SIMD must be in [1, 8]")
        if mode is None:
            mode = "block" if re.search(r"#\s*include\s*[<\"][^>\"]*gtap_block\.cuh", text) else "thread"
        if mode not in ("thread", "block"):
            raise LoweringError(f"unknown mode {mode!r} (thread or block)")
        self.mode = mode
        bs = alldefs.get("GTAP_BLOCK_SIZE")
        if bs is not None:
            ok = ("256",) if mode == "block" else ("64", "128", "192", "256")
            if bs.strip() not in ok:
                raise LoweringError(f"GTAP_BLOCK_SIZE={bs}: workers here are 64-lane wavefronts; the entry launch "
                                    f"takes {' or '.join(ok)} threads per workgroup in {mode} mode")

    def line(self, off: int) -> int:
        return self.text.count("\n", 0, off) + 1

    # ------------------------------------------------------------------ parsing
    def parse(self) -> List[cl.Node]:
        with tempfile.TemporaryDirectory() as d:
            pre = os.path.join(d, "gtap_prelude.h")
            with open(pre, "w", encoding="latin-1") as f:
                f.write(PRELUDE)
            # includes are blanked (same length, so byte offsets stay those of the input)
            blanked = INCLUDE_RE.sub(lambda m: " " * len(m.group(0)), self.text)
            src = os.path.join(d, os.path.basename(self.path) or "input.cu")
            if not src.endswith((".cpp", ".cc", ".cu", ".hip")):
                src += ".cpp"
            with open(src, "w", encoding="latin-1") as f:
                f.write(blanked)
            args = ["-x", "c++", "-std=c++17", "-include", pre, "-Wno-unknown-pragmas", "-Wno-unknown-attributes"]
            args += [f"-D{k}" if v == "" else f"-D{k}={v}" for k, v in self.cli_defines.items()]
            tops, errors = cl.parse(src, args)
        self.errors = errors  # checked against the parts that are lowered or copied
        return tops

    # ------------------------------------------------------------------ analysis
    def collect(self, tops: List[cl.Node]) -> None:
        fn_pragmas = [p for p in self.pragmas if p.kind == "function"]
        funcs = [n for n in tops if n.kind == cl.FUNCTION_DECL and _body(n) is not None]
        self.tasks: Dict[str, TaskFn] = {}
        for p in fn_pragmas:
            cand = [n for n in funcs if n.start >= p.end]
            if not cand:
                raise LoweringError(f"#pragma gtap function at line {self.line(p.start)} is not followed by a function "
                                    "definition")
            n = min(cand, key=lambda x: x.start)
            if not _is_device(n):
                raise LoweringError(f"gtap function {n.spelling} must be __device__")
            t = TaskFn(n.spelling, n, _body(n), n.result_type, n.result_size)
            if t.ret_type != "void" and t.ret_size not in (4, 8):
                raise LoweringError(f"{t.name}: the return type must be 4 or 8 bytes ({t.ret_type})")
            t.taskwaits = [q.start for q in self.pragmas if q.kind == "taskwait" and t.body.start <= q.start < t.body.end]
            self.tasks[t.name] = t
        if not self.tasks:
            raise LoweringError("no #pragma gtap function in the input")
        for i, t in enumerate(sorted(self.tasks.values(), key=lambda x: x.node.start)):
            t.func_id = i + 1
        if self.mode == "block" and any(p.queue != "0" for p in self.pragmas if p.kind in ("task", "taskwait")):
            pass  # block mode has one deque per workgroup: queue() is accepted and ignored, as in the reference
        self.top_nodes = tops
        self.device_vars = [n for n in tops if n.kind == cl.VAR_DECL and (_is_device(n) or _is_constant(n))]
        self.helpers = [n for n in tops if n.kind in (cl.FUNCTION_DECL, FUNCTION_TEMPLATE) and n.spelling not in self.tasks
                        and _is_device(n) and not _is_global(n)]
        self.kernels = [n for n in tops if n.kind == cl.FUNCTION_DECL and _is_global(n) and _body(n) is not None]
        self.types = [n for n in tops if n.kind in (STRUCT_DECL, UNION_DECL, CLASS_DECL, ENUM_DECL, TYPEDEF_DECL,
                                                     TYPE_ALIAS_DECL)]
        self.consts = [n for n in tops if n.kind == cl.VAR_DECL and not (_is_device(n) or _is_constant(n))
                       and re.match(r"\s*(static\s+)?(constexpr|const)\b", self.text[n.start:n.end])]
        for t in self.tasks.values():
            self._vars(t)

    def _is_lds_local(self, n: cl.Node) -> bool:
        """A `__shared__` local (the parse defines the keyword away): the declaration's text
        from the end of the previous statement."""
        k = max(self.text.rfind(c, 0, n.start) for c in ";{}")
        return "__shared__" in self.text[k + 1:n.end]

    def _vars(self, t: TaskFn) -> None:
        names = set()

        def uniq(nm):
            base, k = nm, 1
            while nm in names or nm in ("fr", "lane"):
                k += 1
                nm = f"{base}_{k}"
            names.add(nm)
            return nm

        for c in t.node.children:
            if c.kind == cl.PARM_DECL:
                if c.type_size not in (4, 8):
                    raise LoweringError(f"{t.name}: parameter {c.spelling} must be a 4- or 8-byte scalar or pointer")
                t.params.append(Var(uniq(c.spelling), ("param", c.spelling), c.type, c.type_size, is_param=True))
        for n in t.body.walk():
            if n.kind == cl.VAR_DECL:
                if self._is_lds_local(n):
                    t.lds_locals.add(n.start)  # workgroup memory: stays a declaration, not a frame field
                    continue
                size = n.elem_size if n.array_len >= 0 else n.type_size
                if size not in (4, 8):
                    raise LoweringError(f"{t.name}: local {n.spelling} ({n.type}) at line {self.line(n.start)} must be "
                                        "a 4- or 8-byte scalar or pointer (or an array of them)")
                if n.array_len >= 0 and re.search(r"[={]", self.text[n.start:n.end]):
                    raise LoweringError(f"{t.name}: array local {n.spelling} has an initialiser")
                t.locals.append(Var(uniq(n.spelling), ("local", n.start, n.spelling), n.type, n.type_size,
                                    n.elem_type, n.elem_size, n.array_len))

    # ------------------------------------------------------------------ text rewriting
    def _apply(self, a: int, b: int, reps: List[Tuple[int, int, str]]) -> str:
        """Text [a, b) with the outermost replacements (start, end, text) inside it applied."""
        out, pos = [], a
        for s, e, txt in sorted(reps, key=lambda r: (r[0], -r[1])):
            if s < a or e > b or s < pos:
                continue
            out.append(self.text[pos:s])
            out.append(txt)
            pos = e
        out.append(self.text[pos:b])
        return "".join(out)

    def _shared_lvalue(self, n: cl.Node, fn: cl.Node) -> bool:
        """n (stripped) designates memory other tasks may reach: through a pointer, or a
        file-scope array."""
        n = _strip(n)
        if n.kind == cl.ARRAY_SUBSCRIPT_EXPR:
            base = _strip(n.children[0])
            if _is_array_type(base.type):
                return self._shared_lvalue(base, fn) if base.kind != cl.DECL_REF_EXPR else self._global_ref(base, fn)
            return True
        if n.kind == UNARY_OPERATOR and _unop(self.text, n) == "*":
            return True
        if n.kind == cl.MEMBER_REF_EXPR and n.children:
            if self._is_arrow(n):
                return True
            return self._shared_lvalue(n.children[0], fn)
        if n.kind == cl.DECL_REF_EXPR and _is_array_type(n.type):
            return self._global_ref(n, fn)
        return False

    def _global_ref(self, ref: cl.Node, fn: cl.Node) -> bool:
        if ref.ref is None or ref.ref_kind != cl.VAR_DECL:
            return False
        _, off = ref.ref
        return not (fn.start <= off < fn.end)

    def _is_arrow(self, n: cl.Node) -> bool:
        base = n.children[0]
        return "->" in self.text[base.end:n.end]

    def _wrap_reps(self, root: cl.Node, fn: cl.Node, leaf: List[Tuple[int, int, str]]) -> List[Tuple[int, int, str]]:
        """Coherent loads and stores for the accesses to shared memory under `root`, as
        replacements whose text already contains `leaf` and the nested accesses."""
        if not self.coherent:
            return []
        par = _parents(root)
        cands = []
        for n in root.walk():
            if n.kind in ASM_STMTS:
                raise LoweringError(f"line {self.line(n.start)}: inline asm in task code cannot be made coherent")
            if n.kind == cl.CALL_EXPR and n.spelling in ("memcpy", "memset", "memmove", "__builtin_memcpy",
                                                          "__builtin_memset"):
                raise LoweringError(f"line {self.line(n.start)}: {n.spelling} in task code is not coherent across "
                                    "the GPU's caches; copy with a loop (or gtap_coherent_load/store)")
            if n.kind not in (cl.ARRAY_SUBSCRIPT_EXPR, UNARY_OPERATOR, cl.MEMBER_REF_EXPR):
                continue
            if n.kind == UNARY_OPERATOR and _unop(self.text, n) != "*":
                continue
            if n.kind == cl.MEMBER_REF_EXPR and not n.children:
                continue  # implicit this->member
            if not self._shared_lvalue(n, fn):
                continue
            if n.start >= n.end or n.end > len(self.text):
                continue
            if _is_array_type(n.type) or n.type_size <= 0:
                continue  # an array designator: an address computation, not an access
            ctx = self._access_context(n, par)
            if ctx in ("addr", "base"):
                continue
            cands.append((n, ctx))
        reps: List[Tuple[int, int, str]] = []
        for n, ctx in sorted(cands, key=lambda c: c[0].end - c[0].start):
            inner = [r for r in leaf + reps if n.start <= r[0] and r[1] <= n.end and (r[0], r[1]) != (n.start, n.end)]
            txt = self._apply(n.start, n.end, inner)
            reps.append((n.start, n.end, f"::gtap::coh({txt})" if ctx == "lvalue" else f"::gtap::cld(&({txt}))"))
        return reps

    def _access_context(self, n: cl.Node, par: Dict[int, cl.Node]) -> str:
        child, p = n, par.get(id(n))
        while p is not None and p.kind == PAREN_EXPR:
            child, p = p, par.get(id(p))
        if p is None:
            return "rvalue"
        if p.kind == UNARY_OPERATOR and _unop(self.text, p) == "&":
            return "addr"
        if p.kind == UNARY_OPERATOR and _unop(self.text, p) in ("++", "--"):
            return "lvalue"
        if p.kind == cl.MEMBER_REF_EXPR and p.children and p.children[0] is child and not self._is_arrow(p):
            return "base"
        if p.kind == cl.ARRAY_SUBSCRIPT_EXPR and p.children and p.children[0] is child and _is_array_type(n.type):
            return "base"
        if p.kind in (cl.BINARY_OPERATOR, COMPOUND_ASSIGN) and p.children and p.children[0] is child:
            if p.kind == COMPOUND_ASSIGN or p.spelling == "=":
                return "lvalue"
        return "rvalue"

    # ------------------------------------------------------------------ task functions
    def _var_for_ref(self, t: TaskFn, ref: cl.Node) -> Optional[Var]:
        """The frame field a DeclRefExpr names, if it names a parameter or local of t."""
        if ref.ref is None or ref.ref_kind not in (cl.VAR_DECL, cl.PARM_DECL):
            return None
        f, off = ref.ref
        if not (t.node.start <= off < t.node.end):
            return None
        if ref.ref_kind == cl.PARM_DECL:
            for v in t.params:
                if v.decl == ("param", ref.spelling):
                    return v
            return None
        # a VarDecl's location is its name, its extent starts at the type: the innermost
        # declaration of that name that starts before the location
        cands = [v for v in t.locals if v.decl[2] == ref.spelling and v.decl[1] <= off]
        return max(cands, key=lambda v: v.decl[1]) if cands else None

    def _leaf_reps(self, t: TaskFn) -> List[Tuple[int, int, str]]:
        reps = []
        for n in t.body.walk():
            if n.kind == cl.DECL_REF_EXPR:
                v = self._var_for_ref(t, n)
                if v is not None:
                    reps.append((n.start, n.end, f"fr.{v.name}"))
                elif n.ref_kind == cl.FUNCTION_DECL and n.spelling in self.tasks:
                    reps.append((n.start, n.end, f"{n.spelling}__seq"))
        reps = sorted(reps)
        return reps + self._wrap_reps(t.body, t.node, reps)

    def _seq_reps(self, n: cl.Node) -> List[Tuple[int, int, str]]:
        """Renames of direct task calls plus coherent accesses, for a copied device function."""
        reps = []
        for m in n.walk():
            if m.kind == cl.DECL_REF_EXPR and m.ref_kind == cl.FUNCTION_DECL and m.spelling in self.tasks:
                reps.append((m.start, m.end, f"{m.spelling}__seq"))
        body = _body(n) if n.kind == cl.FUNCTION_DECL else n
        if body is not None:
            reps += self._wrap_reps(body, n, reps)
        return reps

    @staticmethod
    def _int32(type_str: str, size: int) -> bool:
        return size == 4 and type_str.replace("const ", "").strip() in INT32_TYPES

    def _plus_leaves(self, e: cl.Node) -> Optional[List[cl.Node]]:
        """The operands of a tree of `+` over variable references, or None."""
        e = _strip(e)
        if e.kind == cl.BINARY_OPERATOR and e.spelling == "+" and len(e.children) == 2:
            a, b = self._plus_leaves(e.children[0]), self._plus_leaves(e.children[1])
            return None if a is None or b is None else a + b
        return [e] if e.kind == cl.DECL_REF_EXPR else None

    def _sum_join_vars(self, t: TaskFn) -> Optional[set]:
        """The result locals of t when its children's results are only summed: thread mode, one
        taskwait, every task statement `x = f(...);` outside any condition or loop with x a
        distinct 4-byte integer local and f returning a 4-byte integer, and every return after
        the taskwait the sum of exactly those locals, which are used nowhere else. Such a task
        joins additively (each child's result is added into the join word by the atomic that
        counts the join down, with no result slot); the sum of up to 4095 32-bit values is
        exact in the join word and converts to the return type as the C sum does. Any other
        shape keeps result slots."""
        if self.mode == "block" or len(t.taskwaits) != 1 or t.func_id > 127 or not self._int32(t.ret_type, t.ret_size):
            return None
        tw = t.taskwaits[0]
        nodes = list(t.body.walk())
        controls = [n for n in nodes if n.kind in CONTROL_KINDS]
        if any(c.start <= tw < c.end for c in controls):
            return None
        task_pragmas = [p for p in self.pragmas if p.kind == "task" and t.body.start <= p.start < t.body.end]
        if not task_pragmas or len(task_pragmas) > 4095:
            return None
        stmts: Dict[str, cl.Node] = {}
        for p in task_pragmas:
            if p.start > tw:
                return None
            cands = [n for n in nodes if n.start >= p.end and n.kind in (cl.BINARY_OPERATOR, cl.CALL_EXPR, cl.DECL_STMT)]
            if not cands:
                return None
            first = min(cands, key=lambda n: (n.start, -n.end))
            if first.kind != cl.BINARY_OPERATOR or first.spelling != "=" or len(first.children) != 2:
                return None
            if any(c.start <= first.start < c.end for c in controls):
                return None
            lhs, call = _strip(first.children[0]), first.children[1]
            while call.kind != cl.CALL_EXPR and len(call.children) == 1:
                call = call.children[0]
            callee = self.tasks.get(call.spelling) if call.kind == cl.CALL_EXPR else None
            v = self._var_for_ref(t, lhs) if lhs.kind == cl.DECL_REF_EXPR else None
            if (callee is None or v is None or v.is_param or v.array_len >= 0 or v.name in stmts
                    or not self._int32(v.type, v.size) or not self._int32(callee.ret_type, callee.ret_size)):
                return None
            stmts[v.name] = first
        uses = set()
        returns = [n for n in nodes if n.kind == cl.RETURN_STMT and n.start > tw]
        if not returns:
            return None
        for r in returns:
            leaves = self._plus_leaves(r.children[-1]) if r.children else None
            if leaves is None:
                return None
            named = [self._var_for_ref(t, x) for x in leaves]
            if any(x is None for x in named) or sorted(x.name for x in named) != sorted(stmts):
                return None
            uses.update(x.start for x in leaves)
        for n in nodes:
            if n.kind == cl.DECL_REF_EXPR and n.start not in uses:
                v = self._var_for_ref(t, n)
                if v is not None and v.name in stmts and not any(
                        st.children[0].start <= n.start < st.children[0].end for st in stmts.values()):
                    return None
        return set(stmts)

    def lower_task(self, t: TaskFn) -> str:
        t.sum_vars = self._sum_join_vars(t)
        leaf = self._leaf_reps(t)
        stmt: List[Tuple[int, int, str]] = []
        body_pragmas = [p for p in self.pragmas if t.body.start <= p.start < t.body.end]
        nodes = list(t.body.walk())
        state = 0
        for p in body_pragmas:
            if p.kind == "task":
                stmt.append(self._lower_spawn(t, p, nodes, leaf))
            elif p.kind == "taskwait":
                state += 1
                stmt.append((p.start, p.end, self._taskwait_code(t, state, p.queue, leaf)))
            else:
                raise LoweringError(f"{t.name}: #pragma gtap {p.kind} inside a task function (line {self.line(p.start)})")
        # declarations -> assignments to frame fields
        for n in nodes:
            if n.kind == cl.DECL_STMT and not any(s <= n.start and n.end <= e for s, e, _ in stmt):
                if any(v.kind == cl.VAR_DECL and v.start in t.lds_locals for v in n.children):
                    continue  # a __shared__ array keeps its declaration
                parts = []
                for v in n.children:
                    if v.kind != cl.VAR_DECL:
                        raise LoweringError(f"{t.name}: unsupported declaration at line {self.line(n.start)}")
                    var = self._local_by_decl(t, v)
                    init = [] if var.array_len >= 0 else [c for c in v.children if c.kind in EXPR_KINDS]
                    if init:
                        parts.append(f"fr.{var.name} = {self._apply(init[-1].start, init[-1].end, leaf)};")
                end = n.end
                if self.text[end - 1] != ";":
                    end = _stmt_end(self.text, end)
                stmt.append((n.start, end, " ".join(parts) if parts else ";"))
        # returns
        for n in nodes:
            if n.kind == cl.RETURN_STMT:
                end = _stmt_end(self.text, n.end)
                if n.children and t.sum_vars is not None and n.start > t.taskwaits[0]:
                    code = f"{{ __gtap_return(lane, ({t.ret_type})(__gtap_get_join_sum(lane))); return; }}"
                elif n.children:
                    e = n.children[-1]
                    val = self._apply(e.start, e.end, leaf)
                    code = f"{{ __gtap_return(lane, ({t.ret_type})({val})); return; }}"
                else:
                    code = "{ __gtap_finish_task(lane); return; }"
                stmt.append((n.start, end, code))
        stmt.sort()
        for i in range(1, len(stmt)):
            if stmt[i][0] < stmt[i - 1][1]:
                raise LoweringError(f"{t.name}: overlapping constructs at line {self.line(stmt[i][0])}")
        # leaf replacements outside statement-level rewrites
        outer = [r for r in leaf if not any(s <= r[0] and r[1] <= e for s, e, _ in stmt)]
        body = self._apply(t.body.start + 1, t.body.end - 1, stmt + outer)  # inside the braces
        loads = "\n".join(f"    fr.{v.name} = lane.template arg<{v.type}>(offsetof({t.name}__frame, {v.name}));"
                          for v in t.params)
        return f"""template <class L>
__device__ __forceinline__ void {t.name}__task(L& lane) {{
    {t.name}__frame fr;
{loads}
    switch (__gtap_get_task_state(lane)) {{
    case 0: {{
{body}
    }}
    {"" if t.ret_type != "void" else "__gtap_finish_task(lane); return;"}
    default:
        __gtap_finish_task(lane);
        return;
    }}
}}
"""

    def _local_by_decl(self, t: TaskFn, v: cl.Node) -> Var:
        for var in t.locals:
            if var.decl == ("local", v.start, v.spelling):
                return var
        raise LoweringError(f"{t.name}: unknown local {v.spelling}")

    def _lower_spawn(self, t: TaskFn, p: Pragma, nodes: List[cl.Node], leaf) -> Tuple[int, int, str]:
        cands = [n for n in nodes if n.start >= p.end and n.kind in (cl.BINARY_OPERATOR, cl.CALL_EXPR, cl.DECL_STMT)]
        if not cands:
            raise LoweringError(f"{t.name}: #pragma gtap task at line {self.line(p.start)} is not followed by a statement")
        first = min(cands, key=lambda n: (n.start, -n.end))
        lhs_node = None
        if first.kind == cl.BINARY_OPERATOR and first.spelling == "=":
            lhs_node, call = first.children[0], first.children[1]
        elif first.kind == cl.CALL_EXPR:
            call = first
        elif first.kind == cl.DECL_STMT and len(first.children) == 1:
            v = first.children[0]
            init = [c for c in v.children if c.kind in EXPR_KINDS]
            if not init:
                raise LoweringError(f"{t.name}: #pragma gtap task before a declaration without a call "
                                    f"(line {self.line(p.start)})")
            call = init[-1]
            lhs_node = v
        else:
            raise LoweringError(f"{t.name}: #pragma gtap task must precede `lhs = f(args);` or `f(args);` "
                                f"(line {self.line(p.start)})")
        while call.kind != cl.CALL_EXPR and len(call.children) == 1:
            call = call.children[0]
        if call.kind != cl.CALL_EXPR or call.spelling not in self.tasks:
            raise LoweringError(f"{t.name}: the task statement at line {self.line(first.start)} does not call a gtap "
                                "function")
        callee = self.tasks[call.spelling]
        args = [c for c in call.children[1:]]
        if len(args) != len(callee.params):
            raise LoweringError(f"{t.name}: call to {callee.name} with {len(args)} arguments")
        arg_txt = ", ".join(f"({v.type})({self._apply(a.start, a.end, leaf)})" for v, a in zip(callee.params, args))
        end = _stmt_end(self.text, first.end) if self.text[first.end - 1] != ";" else first.end
        q = self._apply_text_queue(p.queue, t)
        joined = any(w > p.start for w in t.taskwaits)
        if not joined:
            if lhs_node is not None:
                raise LoweringError(f"{t.name}: the task at line {self.line(p.start)} assigns a result but no "
                                    "taskwait follows it")
            self.detached_spawns += 1
            code = (f"{{ const {callee.name}__args __a{{{arg_txt}}}; "
                    f"__gtap_spawn_task_detached(lane, F_{callee.name}, (uint32_t)({q}), __a); }}")
            return (p.start, end, code)
        if t.sum_vars is not None:
            return (p.start, end, f"{{ const {callee.name}__args __a{{{arg_txt}}}; "
                                  f"__gtap_spawn_task_sum(lane, F_{callee.name}, (uint32_t)({q}), __a); }}")
        if lhs_node is None:
            off = f"offsetof({t.name}__frame, __discard)"
        else:
            off = self._lhs_offset(t, lhs_node, leaf, callee)
        code = (f"{{ const {callee.name}__args __a{{{arg_txt}}}; const int __off = (int)({off}); "
                f"__gtap_spawn_task(lane, F_{callee.name}, (uint32_t)({q}), __off, __a); }}")
        return (p.start, end, code)

    def _apply_text_queue(self, q: str, t: TaskFn) -> str:
        # queue expressions name parameters or locals of the task function
        names = {v.decl[1] if v.is_param else v.decl[2]: v.name for v in t.params + t.locals}
        return re.sub(r"\b([A-Za-z_]\w*)\b", lambda m: f"fr.{names[m.group(1)]}" if m.group(1) in names else m.group(1), q)

    def _lhs_offset(self, t: TaskFn, lhs: cl.Node, leaf, callee: TaskFn) -> str:
        if lhs.kind == cl.VAR_DECL:
            v = self._local_by_decl(t, lhs)
            self._check_slot(t, v, v.size, callee)
            t.result_targets.add(v.name)
            return f"offsetof({t.name}__frame, {v.name})"
        while lhs.kind not in (cl.DECL_REF_EXPR, cl.ARRAY_SUBSCRIPT_EXPR) and len(lhs.children) == 1:
            lhs = lhs.children[0]
        if lhs.kind == cl.DECL_REF_EXPR:
            v = self._var_for_ref(t, lhs)
            if v is None or v.array_len >= 0:
                raise LoweringError(f"{t.name}: a task's result must go to a local or a local array element")
            self._check_slot(t, v, v.size, callee)
            t.result_targets.add(v.name)
            return f"offsetof({t.name}__frame, {v.name})"
        if lhs.kind == cl.ARRAY_SUBSCRIPT_EXPR:
            base, idx = lhs.children[0], lhs.children[1]
            while base.kind != cl.DECL_REF_EXPR and len(base.children) == 1:
                base = base.children[0]
            v = self._var_for_ref(t, base) if base.kind == cl.DECL_REF_EXPR else None
            if v is None or v.array_len < 0:
                raise LoweringError(f"{t.name}: a task's result must go to a local or a local array element")
            self._check_slot(t, v, v.elem_size, callee)
            t.result_targets.add(v.name)
            return f"offsetof({t.name}__frame, {v.name}) + {v.elem_size} * ({self._apply(idx.start, idx.end, leaf)})"
        raise LoweringError(f"{t.name}: unsupported task result target")

    def _check_slot(self, t: TaskFn, v: Var, size: int, callee: TaskFn) -> None:
        if callee.ret_type == "void" or size != callee.ret_size:
            raise LoweringError(f"{t.name}: result slot {v.name} ({size} B) does not match "
                                f"{callee.name}'s return type ({callee.ret_type})")

    def _taskwait_code(self, t: TaskFn, state: int, queue: str, leaf) -> str:
        q = self._apply_text_queue(queue, t)
        join = "__gtap_set_state_for_join_sum" if t.sum_vars is not None else "__gtap_set_state_for_join"
        return (f"if ({join}(lane, {state}u, (uint32_t)({q}))) {{ {t.name}__save(lane, fr); return; }} "
                f"if (0) {{ case {state}: {t.name}__load(lane, fr); }}")

    # ------------------------------------------------------------------ emission
    def frame_decls(self, t: TaskFn) -> str:
        pad_bytes = (-sum(v.size for v in t.params)) % 8
        params = "\n".join(f"    {v.decl_text()}" for v in t.params)
        pad = f"\n    uint8_t __pad[{pad_bytes}];" if pad_bytes else ""
        loc = "\n".join(f"    {v.decl_text()}" for v in t.locals)
        return f"""struct {t.name}__args {{
{params}{pad}
}};
struct {t.name}__frame {{
{params}{pad}
{loc}
    uint64_t __discard;
}};
static_assert(sizeof({t.name}__args) % 8 == 0, "argument blocks are 8-byte padded");
"""

    def save_load(self, t: TaskFn) -> str:
        # a parameter is saved only if the body assigns it
        assigned = set()
        for n in t.body.walk():
            if n.kind in (cl.BINARY_OPERATOR, COMPOUND_ASSIGN, UNARY_OPERATOR) and n.children:
                tgt = n.children[0]
                while tgt.kind != cl.DECL_REF_EXPR and len(tgt.children) == 1:
                    tgt = tgt.children[0]
                if tgt.kind == cl.DECL_REF_EXPR and (n.kind != cl.BINARY_OPERATOR or n.spelling == "="):
                    v = self._var_for_ref(t, tgt)
                    if v is not None and v.is_param and (n.kind != UNARY_OPERATOR or _unop(self.text, n) in ("++", "--")):
                        assigned.add(v.name)
        sv, ld = [], []
        for v in t.params + t.locals:
            off = f"offsetof({t.name}__frame, {v.name})"
            if t.sum_vars is not None and v.name in t.sum_vars:
                continue  # carried by the additive join, never read
            if v.is_param:
                if v.name in assigned:
                    sv.append(f"    lane.template frame_st<{v.type}>({off}, fr.{v.name});")
                    ld.append(f"    fr.{v.name} = lane.template frame_ld<{v.type}>({off});")
                continue
            if v.array_len >= 0:
                if v.name not in t.result_targets:
                    sv.append(f"    for (int i = 0; i < {v.array_len}; ++i) "
                              f"lane.template frame_st<{v.elem_type}>({off} + {v.elem_size} * i, fr.{v.name}[i]);")
                ld.append(f"    for (int i = 0; i < {v.array_len}; ++i) "
                          f"fr.{v.name}[i] = lane.template frame_ld<{v.elem_type}>({off} + {v.elem_size} * i);")
            else:
                if v.name not in t.result_targets:
                    sv.append(f"    lane.template frame_st<{v.type}>({off}, fr.{v.name});")
                ld.append(f"    fr.{v.name} = lane.template frame_ld<{v.type}>({off});")
        if self.mode == "block":
            # the workgroup's locals are uniform: thread 0 saves them before the task returns
            save = "    if (threadIdx.x == 0) {\n" + "\n".join("    " + s for s in sv) + \
                   "\n        ::gtap::complete_stores();\n    }"
        else:
            save = "\n".join(sv)
        return f"""template <class L>
__device__ __forceinline__ void {t.name}__save(L& lane, const {t.name}__frame& fr) {{
{save}
}}
template <class L>
__device__ __forceinline__ void {t.name}__load(L& lane, {t.name}__frame& fr) {{
{chr(10).join(ld)}
}}
"""

    def copy_function(self, n: cl.Node, rename: Optional[str] = None) -> str:
        """A device function of the input, with direct task calls renamed to their __seq
        copies, its pragmas removed and its shared accesses made coherent."""
        reps = [(p.start, p.end, "") for p in self.pragmas if n.start <= p.start < n.end]
        reps += self._seq_reps(n)
        txt = self._apply(n.start, n.end, reps)
        if rename:
            txt = re.sub(rf"\b{re.escape(rename)}\s*\(", f"{rename}__seq(", txt, count=1)
        if _body(n) is None and n.kind == cl.FUNCTION_DECL:
            txt += ";"
        return txt + "\n"

    def check_errors(self) -> None:
        """Parse errors inside what is lowered or copied are fatal; the rest of the file (host
        code using APIs this parse does not declare) is ignored."""
        spans = [(t.node.start, t.node.end) for t in self.tasks.values()]
        spans += [(n.start, n.end) for n in self.helpers + self.kernels + self.device_vars]
        bad = [m for f, off, m in self.errors if any(a <= off < b for a, b in spans)]
        if bad:
            raise LoweringError("the source does not parse:\n  " + "\n  ".join(bad[:10]))

    def directives(self) -> str:
        """The source's top-level preprocessor lines (definitions and conditionals), except its
        includes and the gtap pragmas, in source order."""
        spans = [(n.start, n.end) for n in self.top_nodes]
        out = []
        for m in DIRECTIVE_RE.finditer(self.text):
            s = m.group(0)
            if INCLUDE_RE.match(s) or PRAGMA_RE.match(s) or re.match(r"\s*#\s*pragma\s+once", s):
                continue
            if any(a <= m.start() < b for a, b in spans):
                continue  # inside a declaration that is copied or dropped whole
            out.append(s.strip())
        return "\n".join(out)

    def std_includes(self) -> str:
        out = []
        for m in INCLUDE_RE.finditer(self.text):
            s = m.group(0).strip()
            if DROP_INCLUDE_RE.search(s):
                continue
            out.append(s)
        return "\n".join(out)

    def entry_kernels(self) -> Tuple[str, str]:
        """The host-launched `__global__` kernels with their entries, and the host helpers."""
        kern_txt, host_txt = [], []
        ns = f"gtap_lowered_{self.name}"
        entries = [p for p in self.pragmas if p.kind == "entry"]
        for k in self.kernels:
            kp = [p for p in entries if k.start <= p.start < k.end]
            if not kp:
                continue  # an ordinary kernel: not carried over (it would run outside the runtime)
            if len(kp) > 1:
                raise LoweringError(f"{k.spelling}: one #pragma gtap entry per kernel")
            p = kp[0]
            nodes = list(k.walk())
            cands = [n for n in nodes if n.start >= p.end and n.kind in (cl.BINARY_OPERATOR, cl.CALL_EXPR)]
            if not cands:
                raise LoweringError(f"{k.spelling}: #pragma gtap entry is not followed by a call")
            first = min(cands, key=lambda n: (n.start, -n.end))
            lhs = None
            if first.kind == cl.BINARY_OPERATOR and first.spelling == "=":
                lhs, call = first.children[0], first.children[1]
            else:
                call = first
            while call.kind != cl.CALL_EXPR and len(call.children) == 1:
                call = call.children[0]
            if call.kind != cl.CALL_EXPR or call.spelling not in self.tasks:
                raise LoweringError(f"{k.spelling}: #pragma gtap entry must call a gtap function")
            t = self.tasks[call.spelling]
            body = _body(k)
            wraps = self._wrap_reps(body, k, [])
            args = ", ".join(f"({v.type})({self._apply(a.start, a.end, wraps)})"
                             for v, a in zip(t.params, call.children[1:]))
            end = _stmt_end(self.text, first.end) if self.text[first.end - 1] != ";" else first.end
            if self.mode == "block":
                call_txt = f"::gtap::entry_block<{ns}::Program>({ns}::gtap_rt_desc, {ns}::F_{t.name}, __gtap_root, __gtap_bits)"
            else:
                q = self._apply(0, 0, []) + (p.queue or "0")
                call_txt = (f"::gtap::entry_thread<{ns}::Program>({ns}::gtap_rt_desc, {ns}::F_{t.name}, "
                            f"(uint32_t)({q}), __gtap_root, __gtap_bits)")
            assign = ""
            if lhs is not None and t.ret_type != "void":
                lhs_txt = self._apply(lhs.start, lhs.end, [r for r in wraps if not (r[0] == lhs.start and r[1] == lhs.end)])
                assign = f" {lhs_txt} = ::gtap::result_as<{t.ret_type}>(__gtap_bits);"
            entry = (f"{{ /* #pragma gtap entry */ const {ns}::{t.name}__args __gtap_root{{{args}}}; int64_t __gtap_bits = 0; "
                     f"if ({call_txt}) {{{assign} }} }}")
            reps = [r for r in wraps if not (p.start <= r[0] < end)] + [(p.start, end, entry)]
            body_txt = self._apply(body.start, body.end, reps)
            params = [c for c in k.children if c.kind == cl.PARM_DECL]
            ptxt = ", ".join(self.text[c.start:c.end] for c in params)
            occ = f" __attribute__((amdgpu_waves_per_eu({self.waves_per_simd})))" if self.waves_per_simd else ""
            kern_txt.append(f"// {k.spelling}: launched by the host with <<<GTAP_GRID_SIZE, GTAP_BLOCK_SIZE>>>\n"
                            f"__global__ __launch_bounds__(GTAP_BLOCK_SIZE){occ} void {k.spelling}({ptxt}) {body_txt}\n")
            # host helpers: the launch as a C function, and the library-launched form
            cparams = ", ".join(f"{c.type} {c.spelling}" for c in params)
            cnames = ", ".join(c.spelling for c in params)
            sep = ", " if cparams else ""
            host_txt.append(f"""// Launches {k.spelling} as a host program does: {k.spelling}<<<grid, block>>>({cnames}).
extern "C" int gtap_launch_{k.spelling}(uint32_t grid, uint32_t block{sep}{cparams}) {{
    {k.spelling}<<<dim3(grid), dim3(block)>>>({cnames});
    return (int)hipGetLastError();
}}
""")
            largs = ", ".join(f"({v.type})({self.text[a.start:a.end]})" for v, a in zip(t.params, call.children[1:]))
            res = "" if t.ret_type == "void" else f"{t.ret_type}* result, "
            copy = "" if t.ret_type == "void" else f"    if (st == 0 && result) memcpy(result, &r, sizeof({t.ret_type}));\n"
            if self.mode == "block":
                run = f"::gtap::host::run_single_block_job<{ns}::Program>({ns}::Program::Args{{0}}, job, &r, stats, {ns}::kTermMode)"
            else:
                run = f"::gtap::host::run_single_job<{ns}::Program>({ns}::Program::Args{{0}}, job, &r, stats, {ns}::kTermMode)"
            host_txt.append(f"""// Runs `{self.text[first.start:first.end]}` of kernel {k.spelling} as one persistent launch
// by the library (the kernel's other statements are not run).
extern "C" int gtap_entry_{k.spelling}({cparams}{sep}{res}gtap_run_stats* stats) {{
    using namespace {ns};
    const {t.name}__args gtap_root_args{{{largs}}};
    gtap_job job = ::gtap::host::make_job(F_{t.name}, (uint32_t)({p.queue or "0"}), &gtap_root_args, sizeof(gtap_root_args));
    int64_t r = 0;
    const int st = {run};
{copy}    return st;
}}
""")
        return "\n".join(kern_txt), "\n".join(host_txt)

    def run(self) -> str:
        tops = self.parse()
        self.collect(tops)
        self.check_errors()
        ns = f"gtap_lowered_{self.name}"
        tasks = sorted(self.tasks.values(), key=lambda x: x.func_id)
        fns = [self.lower_task(t) for t in tasks]  # fills result_targets first
        ids = "\n".join(f"constexpr uint32_t F_{t.name} = {t.func_id};" for t in tasks)
        frames = "\n".join(self.frame_decls(t) for t in tasks)
        seqdecl = "\n".join(self.text[t.node.start:t.body.start].replace(t.name + "(", t.name + "__seq(", 1).strip()
                            + ";" for t in tasks)
        seqs = "\n".join(self.copy_function(t.node, rename=t.name) for t in tasks)
        helpers = "".join(self.copy_function(n) for n in self.helpers)
        saves = "\n".join(self.save_load(t) for t in tasks)
        disp = "\n".join(f"        case F_{t.name}: {t.name}__task(lane); return;" for t in tasks)
        fbytes = ", ".join(f"sizeof({t.name}__frame)" for t in tasks)
        abytes = ", ".join(f"sizeof({t.name}__args)" for t in tasks)
        types = "".join(self.text[n.start:n.end] + ";\n" for n in self.types)
        consts = "".join(self.text[n.start:n.end] + ";\n" for n in self.consts)
        dvars = "".join(self.text[n.start:n.end] + ";\n" for n in self.device_vars)
        kernels, host = self.entry_kernels()
        no_taskwait = not any(t.taskwaits for t in tasks)
        defines = "\n".join(f"#define {k} {v}".rstrip() for k, v in self.cli_defines.items())
        # With every spawn joined, a task ends only after all of its descendants, so the root's
        # end is the moment the last task finishes: the job count ends the launch without the
        # idle-worker accounting that quiescence detection needs.
        term = ("GTAP_TERM_JOBS" if self.first_task_termination or self.detached_spawns == 0
                else "GTAP_TERM_QUIESCE")
        if self.mode == "block":
            prog = f"""struct Program {{{
            chr(10) + f"    static constexpr int kWavesPerSimd = {self.waves_per_simd};" if self.waves_per_simd else ""}
    static constexpr int kFrameWords = (int)(kMaxOf({{{abytes}}}) / 4) < 2 ? 2 : (int)(kMaxOf({{{abytes}}}) / 4);
    static constexpr uint32_t kMaxFunc = {len(tasks)};
    static constexpr uint32_t kFrameBytes = (kMaxOf({{{fbytes}}}) + 15u) & ~15u;
    static_assert(kFrameWords <= 16, "block-mode task arguments are at most 64 bytes");
    struct Args {{
        int32_t unused;
    }};
    struct Shared {{
        int32_t unused;
    }};
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
}};"""
        else:
            prog = f"""struct Program {{
    static constexpr int kQueues = {self.num_queues};{
            chr(10) + f"    static constexpr int kWavesPerSimd = {self.waves_per_simd};" if self.waves_per_simd else ""}
    static constexpr int kFrameWords = (int)(kMaxOf({{{abytes}}}) / 4) < 2 ? 2 : (int)(kMaxOf({{{abytes}}}) / 4);
    static constexpr uint32_t kMaxFunc = {len(tasks)};
    static constexpr uint32_t kFrameBytes = (kMaxOf({{{fbytes}}}) + 15u) & ~15u;{
            chr(10) + "    static constexpr bool kNoTaskwait = true;  // detached spawns only" if no_taskwait else ""}
    static_assert(kFrameWords <= 12, "thread-mode task arguments are at most 48 bytes");
    struct Args {{
        int32_t unused;
    }};
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
}};"""
        return f"""// Generated by gtap_amd.lower from {os.path.basename(self.path)} ({self.mode} mode). Do not edit.
// The source's host functions (main ...) are not carried over: a host program includes or
// links this file and calls gtap_initialize, sets the device variables, and launches the
// entry kernels with <<<GTAP_GRID_SIZE, GTAP_BLOCK_SIZE>>>.
{defines}
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <climits>
#include <cmath>
#include <initializer_list>

#include "device/gtap_lowered.hpp"
#include "host/gtap_launch.hpp"
{self.std_includes()}

template <class T>
__device__ __forceinline__ T gtap_coherent_load(const T* p) {{ return ::gtap::cld(p); }}
template <class T>
__device__ __forceinline__ void gtap_coherent_store(T* p, T v) {{ ::gtap::cst(p, v); }}

// ---- the source's definitions, types and device variables
{self.directives()}
{types}{consts}{dvars}
// ---- device functions of the source (shared accesses made coherent)
{seqdecl}
{helpers}
{seqs}
// ---- task functions as state machines over this runtime's device ABI
namespace {ns} {{
using namespace ::gtap;
{ids}
constexpr uint32_t kTermMode = {term};

{frames}
{saves}
constexpr uint32_t kMaxOf(std::initializer_list<size_t> l) {{
    size_t m = 0;
    for (size_t x : l) m = x > m ? x : m;
    return (uint32_t)m;
}}

{prog}

{chr(10).join(fns)}
template <class L>
__device__ __forceinline__ void Program::run(L& lane) {{
    switch (lane.func) {{
{disp}
        default: __gtap_finish_task(lane); return;
    }}
}}

// The runtime descriptor of the host-launched entry kernels, pushed by gtap_initialize and
// gtap_reset through the setup function this module registers when it is loaded.
__constant__ ::gtap::GtapDev gtap_rt_desc;
static int gtap_module_setup(const void* desc, uint64_t bytes) {{
    if (bytes != sizeof(::gtap::GtapDev)) return (int)hipErrorInvalidValue;
    ::gtap::GtapDev d;
    memcpy(&d, desc, sizeof(d));
    d.term_mode = kTermMode;
    return (int)hipMemcpyToSymbol(HIP_SYMBOL(gtap_rt_desc), &d, sizeof(d));
}}
static const int gtap_module_registered = gtap_register_entry_module(&gtap_module_setup);

}}  // namespace {ns}

// ---- entry kernels
{kernels}
// ---- host helpers
// Bytes per task record this program needs (gtap_config.record_stride must be at least this).
extern "C" uint32_t gtap_lowered_{self.name}_record_bytes(void) {{
    return {ns}::Program::kFrameBytes + ::gtap::kHdrBytes;
}}

// The co-resident launch of this module's entry kernels: every workgroup that fits on the
// GPU at once (256 threads each), at most the runtime's provisioned workers.
extern "C" int gtap_lowered_{self.name}_launch_dims(uint32_t* grid, uint32_t* block) {{
    int dev = 0, nb = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return (int)hipErrorNoDevice;
    uint32_t best = 0;
{self._dims_body()}
    const uint32_t cap = (uint32_t)prop.multiProcessorCount * {8 if self.mode == "thread" else 32};
    if (best == 0 || best > cap) best = cap;
    if (grid) *grid = best;
    if (block) *block = GTAP_BLOCK_SIZE;
    return 0;
}}
// The host symbol of a device variable of the source, by name (for hipMemcpyToSymbol /
// hipMemcpyFromSymbol from callers that do not compile against this file), and its size.
extern "C" const void* gtap_lowered_{self.name}_symbol(const char* name, uint64_t* bytes) {{
{self._symbol_table()}
    return nullptr;
}}
static inline unsigned gtap_lowered_{self.name}_grid_size() {{
    uint32_t g = 0;
    return gtap_lowered_{self.name}_launch_dims(&g, nullptr) == 0 ? g : 1u;
}}
#ifndef GTAP_GRID_SIZE
#define GTAP_GRID_SIZE (gtap_lowered_{self.name}_grid_size())
#endif

{host}"""

    def _symbol_table(self) -> str:
        out = []
        for n in self.device_vars:
            v = n.spelling
            out.append(f'    if (strcmp(name, "{v}") == 0) {{\n'
                       f'        if (bytes) *bytes = sizeof({v});\n'
                       f'        return reinterpret_cast<const void*>(&{v});\n    }}')
        return "\n".join(out)

    def _dims_body(self) -> str:
        out = []
        for k in self.kernels:
            if not any(p.kind == "entry" and k.start <= p.start < k.end for p in self.pragmas):
                continue
            out.append(f"    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&{k.spelling}), "
                       f"GTAP_BLOCK_SIZE, 0) == hipSuccess && nb > 0) {{\n"
                       f"        const uint32_t g = (uint32_t)nb * (uint32_t)prop.multiProcessorCount;\n"
                       f"        if (best == 0 || g < best) best = g;\n    }}")
        return "\n".join(out)


def lower(text: str, path: str = "input.cu", name: Optional[str] = None, defines: Optional[Dict[str, str]] = None,
          mode: Optional[str] = None, coherent: bool = True) -> str:
    """Lowers one `#pragma gtap` source; returns the generated HIP source. `defines` are the
    compile-time switches of the reference's command line (GTAP_NUM_QUEUES, ...); `mode` is
    "thread" or "block" (default: block when the source includes gtap_block.cuh)."""
    if name is None:
        name = os.path.basename(path).split(".")[0]
    return _Lowerer(text, path, name, defines, mode, coherent).run().encode("latin-1").decode("utf-8")


def lower_file(src: str, out: str, name: Optional[str] = None, defines: Optional[Dict[str, str]] = None,
               mode: Optional[str] = None, coherent: bool = True) -> None:
    with open(src, encoding="utf-8") as f:
        text = f.read()
    code = lower(text, src, name, defines, mode, coherent)
    with open(out + ".tmp", "w", encoding="utf-8") as f:
        f.write(code)
    os.replace(out + ".tmp", out)
