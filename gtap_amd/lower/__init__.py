"""gtap_amd.lower -- source-to-source lowering of `#pragma gtap` task programs to HIP for this
runtime (the role of the reference's clang-gtap, which is not part of the reference tree;
its contract is the device ABI of runtime/detail/thread/gtap_runtime_thread.cuh:663-847).

Input: a CUDA/HIP source in the reference's dialect --

    #pragma gtap function
    __device__ long fib(int n) {
        if (n < 2) return n;
        long a, b;
        #pragma gtap task
        a = fib(n - 1);
        #pragma gtap task [queue(expr)]
        b = fib(n - 2);
        #pragma gtap taskwait [queue(expr)]
        return a + b;
    }
    __global__ void kernel(int n) {
        #pragma gtap entry
        d_result = fib(n);
    }

Output: one .hip file with, per task function f,
  * `f__args` (the parameters, padded to 8 bytes) and `f__frame` (parameters, then every
    local of f, then a discard slot): the task record's frame;
  * `f__task(lane)`: a switch over the resume state in which every local lives in the
    frame struct `fr` (so control may re-enter the body at a taskwait, protothread style);
    a spawn names the child's function id and the byte offset of the result slot (`lhs`)
    inside the parent's frame; a taskwait that suspends saves the frame (result slots
    excepted), and the resumed task reloads it;
  * `f__seq`: an ordinary device function for direct (non-task) calls;
and one Program struct (function-id dispatch) plus, per `#pragma gtap entry` kernel, an
`extern "C" int gtap_entry_<kernel>(<kernel params>, <result type>* result,
gtap_run_stats* stats)` that runs the root as one persistent launch.

Restrictions (reported as errors): thread mode only; parameters and locals are 4- or 8-byte
scalars, pointers, or arrays of them, without array initialisers; a task's `lhs` is a local
or an element `local[i]` of a local array whose type has the callee's return size. Data that
tasks share through pointers must be read and written with gtap_coherent_load /
gtap_coherent_store (device-scope coherent accesses across the GPU's XCD caches).
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
NUMQ_RE = re.compile(r"^[ \t]*#[ \t]*define[ \t]+GTAP_NUM_QUEUES[ \t]+(\d+)", re.M)

PRELUDE = """#define __device__ __attribute__((annotate("gtap_device")))
#define __global__ __attribute__((annotate("gtap_global")))
#define __host__
#define __shared__
#define __forceinline__ inline
#define __restrict__
#define __launch_bounds__(...)
typedef signed char int8_t; typedef unsigned char uint8_t; typedef short int16_t;
typedef unsigned short uint16_t; typedef int int32_t; typedef unsigned int uint32_t;
typedef long int64_t; typedef unsigned long uint64_t; typedef unsigned long size_t;
typedef long ptrdiff_t; typedef long intptr_t; typedef unsigned long uintptr_t;
#define NULL 0
struct __gtap_dim3 { unsigned x, y, z; };
extern const __gtap_dim3 threadIdx, blockIdx, blockDim, gridDim;
extern const int warpSize;
void __syncthreads();
void __threadfence();
int atomicAdd(int*, int); unsigned atomicAdd(unsigned*, unsigned);
unsigned long long atomicAdd(unsigned long long*, unsigned long long); float atomicAdd(float*, float);
int atomicMin(int*, int); unsigned atomicMin(unsigned*, unsigned);
int atomicMax(int*, int); unsigned atomicMax(unsigned*, unsigned);
int atomicCAS(int*, int, int); unsigned atomicCAS(unsigned*, unsigned, unsigned);
int atomicExch(int*, int); unsigned atomicExch(unsigned*, unsigned);
template <class T> T gtap_coherent_load(const T* p);
template <class T> void gtap_coherent_store(T* p, T v);
"""

ANNOTATE_ATTR = 406
UNARY_OPERATOR = 112
COMPOUND_ASSIGN = 115


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
    decl: Tuple     # (file, offset) identity of the declaration
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


def _stmt_end(text: str, pos: int) -> int:
    """Offset just past the ';' that ends the expression statement reaching `pos`."""
    i = pos
    while i < len(text) and text[i] in " \t\r\n":
        i += 1
    if i >= len(text) or text[i] != ";":
        raise LoweringError(f"expected ';' at offset {pos}")
    return i + 1


def _is_device(n: cl.Node) -> bool:
    return any(c.kind == ANNOTATE_ATTR and c.spelling == "gtap_device" for c in n.children)


def _is_global(n: cl.Node) -> bool:
    return any(c.kind == ANNOTATE_ATTR and c.spelling == "gtap_global" for c in n.children)


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


class _Lowerer:
    def __init__(self, text: str, path: str, name: str):
        self.text = text
        self.path = path
        self.name = re.sub(r"\W", "_", name)
        self.pragmas: List[Pragma] = []
        for m in PRAGMA_RE.finditer(text):
            rest = m.group(2).strip()
            q = QUEUE_RE.match(rest)
            self.pragmas.append(Pragma(m.group(1), m.start(), m.end(), q.group(1).strip() if q else "0"))
        nq = NUMQ_RE.search(text)
        self.num_queues = int(nq.group(1)) if nq else 1

    # ------------------------------------------------------------------ parsing
    def parse(self) -> List[cl.Node]:
        with tempfile.TemporaryDirectory() as d:
            pre = os.path.join(d, "gtap_prelude.h")
            with open(pre, "w") as f:
                f.write(PRELUDE)
            # includes are blanked (same length, so byte offsets stay those of the input)
            blanked = INCLUDE_RE.sub(lambda m: " " * len(m.group(0)), self.text)
            src = os.path.join(d, os.path.basename(self.path) or "input.cu")
            if not src.endswith((".cpp", ".cc", ".cu", ".hip")):
                src += ".cpp"
            with open(src, "w") as f:
                f.write(blanked)
            tops, errors = cl.parse(src, ["-x", "c++", "-std=c++17", "-include", pre, "-Wno-unknown-pragmas",
                                          "-Wno-unknown-attributes"])
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
                raise LoweringError(f"#pragma gtap function at offset {p.start} is not followed by a function definition")
            n = min(cand, key=lambda x: x.start)
            if not _is_device(n):
                raise LoweringError(f"gtap function {n.spelling} must be __device__")
            t = TaskFn(n.spelling, n, _body(n), n.result_type, n.result_size)
            if t.ret_type != "void" and t.ret_size not in (4, 8):
                raise LoweringError(f"{t.name}: the return type must be 4 or 8 bytes ({t.ret_type})")
            self.tasks[t.name] = t
        if not self.tasks:
            raise LoweringError("no #pragma gtap function in the input")
        for i, t in enumerate(sorted(self.tasks.values(), key=lambda x: x.node.start)):
            t.func_id = i + 1
        self.others = [n for n in tops if n.spelling not in self.tasks or n.kind != cl.FUNCTION_DECL]
        for t in self.tasks.values():
            self._vars(t)

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
                size = n.elem_size if n.array_len >= 0 else n.type_size
                if size not in (4, 8):
                    raise LoweringError(f"{t.name}: local {n.spelling} ({n.type}) must be 4- or 8-byte scalars")
                if n.array_len >= 0 and re.search(r"[={]", self.text[n.start:n.end]):
                    raise LoweringError(f"{t.name}: array local {n.spelling} has an initialiser")
                t.locals.append(Var(uniq(n.spelling), ("local", n.start, n.spelling), n.type, n.type_size,
                                    n.elem_type, n.elem_size, n.array_len))

    # ------------------------------------------------------------------ rewriting
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
        return sorted(reps)

    def _apply(self, a: int, b: int, reps: List[Tuple[int, int, str]]) -> str:
        out, pos = [], a
        for s, e, txt in reps:
            if s < a or e > b or s < pos:
                continue
            out.append(self.text[pos:s])
            out.append(txt)
            pos = e
        out.append(self.text[pos:b])
        return "".join(out)

    def lower_task(self, t: TaskFn) -> str:
        par = _parents(t.body)
        leaf = self._leaf_reps(t)
        stmt: List[Tuple[int, int, str]] = []
        body_pragmas = [p for p in self.pragmas if t.body.start <= p.start < t.body.end]
        nodes = list(t.body.walk())
        state = 0
        pending_targets = set()
        for p in body_pragmas:
            if p.kind == "task":
                stmt.append(self._lower_spawn(t, p, nodes, leaf))
            elif p.kind == "taskwait":
                state += 1
                stmt.append((p.start, p.end, self._taskwait_code(t, state, p.queue, leaf)))
            else:
                raise LoweringError(f"{t.name}: #pragma gtap {p.kind} inside a task function")
        # declarations -> assignments to frame fields
        for n in nodes:
            if n.kind == cl.DECL_STMT and not any(s <= n.start and n.end <= e for s, e, _ in stmt):
                parts = []
                for v in n.children:
                    if v.kind != cl.VAR_DECL:
                        raise LoweringError(f"{t.name}: unsupported declaration at offset {n.start}")
                    var = self._local_by_decl(t, v)
                    init = [] if var.array_len >= 0 else [c for c in v.children if 100 <= c.kind < 200]
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
                if n.children:
                    e = n.children[-1]
                    val = self._apply(e.start, e.end, leaf)
                    code = f"{{ __gtap_return(lane, ({t.ret_type})({val})); return; }}"
                else:
                    code = "{ __gtap_finish_task(lane); return; }"
                stmt.append((n.start, end, code))
        stmt.sort()
        for i in range(1, len(stmt)):
            if stmt[i][0] < stmt[i - 1][1]:
                raise LoweringError(f"{t.name}: overlapping constructs at offset {stmt[i][0]}")
        # leaf replacements outside statement-level rewrites
        outer = [r for r in leaf if not any(s <= r[0] and r[1] <= e for s, e, _ in stmt)]
        reps = sorted(stmt + outer)
        body = self._apply(t.body.start + 1, t.body.end - 1, reps)  # inside the braces
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
            raise LoweringError(f"{t.name}: #pragma gtap task is not followed by a statement")
        first = min(cands, key=lambda n: (n.start, -n.end))
        lhs_node = None
        if first.kind == cl.BINARY_OPERATOR and first.spelling == "=":
            lhs_node, call = first.children[0], first.children[1]
        elif first.kind == cl.CALL_EXPR:
            call = first
        elif first.kind == cl.DECL_STMT and len(first.children) == 1:
            v = first.children[0]
            init = [c for c in v.children if 100 <= c.kind < 200]
            if not init:
                raise LoweringError(f"{t.name}: #pragma gtap task before a declaration without a call")
            call = init[-1]
            lhs_node = v
        else:
            raise LoweringError(f"{t.name}: #pragma gtap task must precede `lhs = f(args);` or `f(args);`")
        while call.kind != cl.CALL_EXPR and len(call.children) == 1:
            call = call.children[0]
        if call.kind != cl.CALL_EXPR or call.spelling not in self.tasks:
            raise LoweringError(f"{t.name}: the task statement at offset {first.start} does not call a gtap function")
        callee = self.tasks[call.spelling]
        args = [c for c in call.children[1:]]
        if len(args) != len(callee.params):
            raise LoweringError(f"{t.name}: call to {callee.name} with {len(args)} arguments")
        arg_txt = ", ".join(f"({v.type})({self._apply(a.start, a.end, leaf)})" for v, a in zip(callee.params, args))
        end = _stmt_end(self.text, first.end) if self.text[first.end - 1] != ";" else first.end
        q = self._apply_text_queue(p.queue, t)
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
        return (f"if (__gtap_set_state_for_join(lane, {state}u, (uint32_t)({q}))) {{ {t.name}__save(lane, fr); return; }} "
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
                    if v is not None and v.is_param and (n.kind != UNARY_OPERATOR or n.spelling in ("++", "--")):
                        assigned.add(v.name)
        sv, ld = [], []
        for v in t.params + t.locals:
            off = f"offsetof({t.name}__frame, {v.name})"
            if v.is_param:
                if v.name in assigned:
                    sv.append(f"    lane.template frame_st<{v.type}>({off}, fr.{v.name});")
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
        return f"""template <class L>
__device__ __forceinline__ void {t.name}__save(L& lane, const {t.name}__frame& fr) {{
{chr(10).join(sv)}
}}
template <class L>
__device__ __forceinline__ void {t.name}__load(L& lane, {t.name}__frame& fr) {{
{chr(10).join(ld)}
}}
"""

    def seq_copy(self, t: TaskFn) -> str:
        n = t.node
        reps = []
        for p in self.pragmas:
            if n.start <= p.start < n.end:
                reps.append((p.start, p.end, ""))
        for m in n.walk():
            if m.kind == cl.DECL_REF_EXPR and m.ref_kind == cl.FUNCTION_DECL and m.spelling in self.tasks:
                reps.append((m.start, m.end, f"{m.spelling}__seq"))
        # the function's own name in its declarator
        txt = self._apply(n.start, n.end, sorted(reps))
        txt = re.sub(rf"\b{re.escape(t.name)}\s*\(", f"{t.name}__seq(", txt, count=1)
        return txt + "\n"

    def entries(self, tops: List[cl.Node]) -> str:
        out = []
        for p in (p for p in self.pragmas if p.kind == "entry"):
            kern = [n for n in tops if n.kind == cl.FUNCTION_DECL and n.start <= p.start < n.end]
            if not kern or not _is_global(kern[0]):
                raise LoweringError("#pragma gtap entry must be inside a __global__ function")
            k = kern[0]
            nodes = list(k.walk())
            cands = [n for n in nodes if n.start >= p.end and n.kind in (cl.BINARY_OPERATOR, cl.CALL_EXPR)]
            first = min(cands, key=lambda n: (n.start, -n.end))
            call = first.children[1] if first.kind == cl.BINARY_OPERATOR else first
            while call.kind != cl.CALL_EXPR and len(call.children) == 1:
                call = call.children[0]
            if call.spelling not in self.tasks:
                raise LoweringError(f"{k.spelling}: #pragma gtap entry must call a gtap function")
            t = self.tasks[call.spelling]
            kparams = ", ".join(f"{c.type} {c.spelling}" for c in k.children if c.kind == cl.PARM_DECL)
            args = ", ".join(f"({v.type})({self.text[a.start:a.end]})" for v, a in zip(t.params, call.children[1:]))
            q = p.queue
            res = "" if t.ret_type == "void" else f"{t.ret_type}* result, "
            copy = "" if t.ret_type == "void" else f"    if (st == 0 && result) memcpy(result, &r, sizeof({t.ret_type}));\n"
            sep = ", " if kparams else ""
            out.append(f"""// Runs `{self.text[first.start:first.end]}` of kernel {k.spelling} as one persistent launch.
extern "C" int gtap_entry_{k.spelling}({kparams}{sep}{res}gtap_run_stats* stats) {{
    using namespace gtap_lowered_{self.name};
    const {t.name}__args gtap_root_args{{{args}}};
    gtap_job job = gtap::host::make_job(F_{t.name}, (uint32_t)({q}), &gtap_root_args, sizeof(gtap_root_args));
    int64_t r = 0;
    const int st = gtap::host::run_single_job<Program>(Program::Args{{0}}, job, &r, stats, GTAP_TERM_JOBS);
{copy}    return st;
}}
""")
        return "\n".join(out)

    def others_text(self) -> str:
        """Device helpers, types and constants of the input that the tasks may use."""
        out = []
        for n in self.others:
            txt = self.text[n.start:n.end]
            if n.kind == cl.FUNCTION_DECL:
                if not _is_device(n) or _is_global(n) or n.spelling == "main" or n.spelling in self.tasks:
                    continue
                out.append(txt + "\n")
            elif n.kind == cl.VAR_DECL:
                if _is_device(n):
                    continue  # device globals written by an entry kernel's lhs are replaced
                out.append(txt + ";\n")
            elif n.kind in (2, 3, 4, 5, 20, 36):  # struct, union, class, enum, typedef, alias
                out.append(txt + ";\n")
        return "".join(out)

    def check_errors(self, tops: List[cl.Node]) -> None:
        """Parse errors inside what is lowered or copied are fatal; the rest of the file (host
        code using APIs this parse does not declare) is ignored."""
        spans = [(t.node.start, t.node.end) for t in self.tasks.values()]
        for n in tops:
            if n.kind == cl.FUNCTION_DECL and (_is_global(n) or (_is_device(n) and n.spelling not in self.tasks)):
                spans.append((n.start, n.end))
        bad = [m for f, off, m in self.errors if any(a <= off < b for a, b in spans)]
        if bad:
            raise LoweringError("the source does not parse:\n  " + "\n  ".join(bad[:10]))

    def run(self) -> str:
        tops = self.parse()
        self.collect(tops)
        self.check_errors(tops)
        tasks = sorted(self.tasks.values(), key=lambda x: x.func_id)
        fns = [self.lower_task(t) for t in tasks]  # fills result_targets first
        ids = "\n".join(f"constexpr uint32_t F_{t.name} = {t.func_id};" for t in tasks)
        frames = "\n".join(self.frame_decls(t) for t in tasks)
        seqdecl = "\n".join(self.text[t.node.start:t.body.start].replace(t.name + "(", t.name + "__seq(", 1).strip()
                            + ";" for t in tasks)
        seqs = "\n".join(self.seq_copy(t) for t in tasks)
        saves = "\n".join(self.save_load(t) for t in tasks)
        kfw = max((sum(v.size for v in t.params) + 7) // 8 * 2 for t in tasks)
        kfw = max(kfw, 2)
        disp = "\n".join(f"        case F_{t.name}: {t.name}__task(lane); return;" for t in tasks)
        fbytes = " , ".join(f"sizeof({t.name}__frame)" for t in tasks)
        return f"""// Generated by gtap_amd.lower from {os.path.basename(self.path)}. Do not edit.
#include <cstddef>
#include <cstring>
#include <initializer_list>

#include "device/gtap_abi.hpp"
#include "host/gtap_launch.hpp"

template <class T>
__device__ __forceinline__ T gtap_coherent_load(const T* p) {{ return gtap::ldT<T>(p); }}
template <class T>
__device__ __forceinline__ void gtap_coherent_store(T* p, T v) {{ gtap::stT<T>(p, v); }}

namespace gtap_lowered_{self.name} {{
using namespace gtap;
{ids}

{self.others_text()}
{seqdecl}

{frames}
{seqs}
{saves}
constexpr uint32_t kMaxFrame(std::initializer_list<size_t> l) {{
    size_t m = 0;
    for (size_t x : l) m = x > m ? x : m;
    return (uint32_t)m;
}}

struct Program {{
    static constexpr int kQueues = {self.num_queues};
    static constexpr int kFrameWords = {kfw};
    static constexpr uint32_t kMaxFunc = {len(tasks)};
    static constexpr uint32_t kFrameBytes = (kMaxFrame({{{fbytes}}}) + 15u) & ~15u;
    struct Args {{
        int32_t unused;
    }};
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
}};

{chr(10).join(fns)}
template <class L>
__device__ __forceinline__ void Program::run(L& lane) {{
    switch (lane.func) {{
{disp}
        default: __gtap_finish_task(lane); return;
    }}
}}

}}  // namespace gtap_lowered_{self.name}

// Bytes per task record this program needs (gtap_config.record_stride must be at least this).
extern "C" uint32_t gtap_lowered_{self.name}_record_bytes(void) {{
    return gtap_lowered_{self.name}::Program::kFrameBytes + gtap::kHdrBytes;
}}

{self.entries(tops)}
"""


def lower(text: str, path: str = "input.cu", name: Optional[str] = None) -> str:
    """Lowers one `#pragma gtap` source; returns the generated HIP source."""
    if name is None:
        name = os.path.basename(path).split(".")[0]
    return _Lowerer(text, path, name).run()


def lower_file(src: str, out: str, name: Optional[str] = None) -> None:
    with open(src) as f:
        text = f.read()
    code = lower(text, src, name)
    with open(out + ".tmp", "w") as f:
        f.write(code)
    os.replace(out + ".tmp", out)
