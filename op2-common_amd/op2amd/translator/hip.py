"""HIP (gfx950) code generation for translated OP2 loops: one kernel file per
program holding, per loop, a device copy of the user kernel, a HIP kernel and
the op_par_loop stub that drives it through the runtime ABI
(include/op2_hip_rt.h); plus the program's constants and their hook.

What a stub does, per loop (the reference's HIP target does the same job with
its own runtime, translator-v2/resources/templates/hip/loop_host.hpp.jinja):
- plans: loops whose increments all go through one map into one dat, and
  whose other side effects are idempotent, run the runtime's owner partition
  by default (each block evaluates the elements touching its targets once,
  increments go to LDS, each owned target sums them in element order: no
  atomics, deterministic, the sequential executor's order); atomics (the
  reference's `atomics` option) and a flat colouring, one launch per colour,
  are selectable; loops with indirect OP_RW/OP_WRITE args run coloured;
- dats in SoA layout are copied into registers around the user kernel call
  (the user kernel always sees AoS pointers, no index rewriting);
- gbl OP_READ values travel by value in the kernel arguments; OP_INC/OP_MIN/
  OP_MAX reduce per workgroup into partials and a one-workgroup finish kernel
  writes the loop's result to the runtime's result slot; a linked
  op_arg_info reports the payload of the first element (lowest index)
  reaching the extremum, as the sequential executor does.
"""

import re

from . import source
from .op import ArgDat, ArgGbl, ArgIdx, ArgInfo

PLAN_ATOMIC, PLAN_COLOR, PLAN_OWNER = 1, 2, 5


def _ctype(t):
    return t


def _size(t):
    return {"double": 8, "float": 4, "int": 4, "unsigned int": 4, "long long": 8,
            "bool": 1, "char": 1}[t]


def _pad8(n):
    return (n + 7) // 8 * 8


def _device_kernel(text):
    return "__device__ " + text.lstrip()


def _arith_estimate(kernel_text):
    """Multiply-add equivalents per element of a user kernel, for the cost
    model: its arithmetic operators, a division or a call of sqrt/pow/exp/
    log/trigonometry counted as ten, an operator inside a `for` loop's body
    as four (loops over components are short); clamped to 4-64."""
    body = source.strip_comments(kernel_text[kernel_text.index("{"):])
    body = re.sub(r"->|\+\+|--", " ", body)

    def ops(t):
        return len(re.findall(r"[*+\-]", t)) + 10 * (
            len(re.findall(r"/", t)) +
            len(re.findall(r"\b(?:sqrt|pow|exp|log|sin|cos|tan|atan2?)\w*\s*\(", t)))
    total = ops(body)
    for m in re.finditer(r"\bfor\s*\(", body):
        close = source.match_close(body, m.end() - 1)
        k = close + 1
        while k < len(body) and body[k] in " \t\r\n":
            k += 1
        inner = body[k:source.match_close(body, k) + 1] if k < len(body) and body[k] == "{" \
            else body[k:body.find(";", k) + 1]
        total += 3 * ops(inner)
    return max(4, min(64, total))


class LoopGen:
    def __init__(self, loop, ns, arith=16):
        self.L, self.ns, self.arith = loop, ns, arith
        self.info_of = {}  # gbl arg index -> info arg index
        for j, a in enumerate(loop.args):
            if isinstance(a, ArgInfo):
                if a.ref in self.info_of:
                    raise ValueError(f"{loop.kernel}: two op_arg_info args for arg {a.ref}")
                self.info_of[a.ref] = j
        self.reds = [j for j, a in enumerate(loop.args)
                     if isinstance(a, ArgGbl) and a.acc in ("OP_INC", "OP_MIN", "OP_MAX")]
        for j in self.reds:
            if j in self.info_of and loop.args[j].dim != 1:
                raise ValueError(f"{loop.kernel}: op_arg_info needs a gbl of dim 1")
        for j, a in enumerate(loop.args):
            if isinstance(a, ArgGbl) and a.acc in ("OP_WRITE", "OP_RW"):
                raise ValueError(f"{loop.kernel}: gbl OP_WRITE/OP_RW is not supported")
        # partial record layout
        self.rec_off, off = {}, 0
        for j in self.reds:
            a = loop.args[j]
            self.rec_off[j] = off
            off += _pad8(a.dim * _size(a.typ))
            if j in self.info_of:
                ia = loop.args[self.info_of[j]]
                off += _pad8(ia.dim * _size(ia.typ)) + 8
        self.rec = max(off, 8)
        self.owner = self._owner_eligible()
        # args on one map read its rows through the first of them, so the
        # compiler sees one base pointer and can merge the loads of a row
        # (optional args keep their own, which is null when they are off)
        self.rep, first = {}, {}
        for j, a in enumerate(loop.args):
            mapped = (isinstance(a, ArgDat) and not a.direct) or (isinstance(a, ArgIdx) and a.map)
            if mapped and getattr(a, "opt", None) is None:
                self.rep[j] = first.setdefault(a.map, j)
        if loop.indirect_writes():
            self.def_plan, self.mask = PLAN_COLOR, 1 << PLAN_COLOR
        elif loop.indirect_incs():
            self.def_plan, self.mask = PLAN_ATOMIC, (1 << PLAN_ATOMIC) | (1 << PLAN_COLOR)
            if self.owner:
                # owner-compute: no atomics, deterministic sums in element order
                self.def_plan = PLAN_OWNER
                self.mask |= 1 << PLAN_OWNER
        else:
            self.def_plan, self.mask = PLAN_ATOMIC, 1 << PLAN_ATOMIC

    def _owner_eligible(self):
        """The owner partition evaluates elements on a block boundary twice, so
        every side effect besides the increments must be idempotent: only
        OP_READ/OP_WRITE direct dats, indirect reads, gbl reads and indices,
        and increments all through one map into one dat."""
        incs = self.L.indirect_incs()
        if not incs or self.L.indirect_writes() or self.reds:
            return False
        if len({(a.dat, a.map) for a in incs}) != 1:
            return False
        return all(a.acc in ("OP_READ", "OP_WRITE") for a in self.L.args
                   if isinstance(a, ArgDat) and a.direct)

    @property
    def name(self):
        return f"{self.ns}_{self.L.index}"

    # ---- kernel argument struct ----
    def args_struct(self):
        f = []
        for j, a in enumerate(self.L.args):
            if isinstance(a, ArgDat):
                f.append(f"  {a.typ} *d{j}; int s{j}; int opt{j};")
                if not a.direct:
                    f.append(f"  const int *m{j}; int md{j}; const int *r{j};")
                    f.append(f"  const unsigned short *h{j}; const int *hb{j};")
            elif isinstance(a, ArgGbl) and a.acc == "OP_READ":
                f.append(f"  {a.typ} g{j}[{a.dim}]; int opt{j};")
            elif isinstance(a, ArgIdx) and a.map:
                f.append(f"  const int *m{j}; int md{j}; const int *r{j};")
                f.append(f"  const unsigned short *h{j}; const int *hb{j};")
            if isinstance(a, ArgIdx):
                # the application's index of each element of the indexed set
                # under transparent renumbering (OP_HIP_RENUMBER), else null
                f.append(f"  const int *oi{j};")
        f.append("  const int *on;  // the loop set's application indices (OP_HIP_RENUMBER), or null")
        f.append("  int _unused;")
        return f"struct {self.name}_args {{\n" + "\n".join(f) + "\n};\n"

    def _staged(self, j, a):
        """AoS data of arg j goes through registers when the arg reads it
        (OP_READ, OP_RW, a direct OP_INC) and no other arg of the loop writes
        its dat: the kernel's loads then issue together, not each after the
        previous store through a pointer that might alias it.  OP_WRITE args
        keep pointers into the dat, so components the kernel leaves alone
        stay as they were."""
        if a.acc == "OP_WRITE":
            return False
        for i, b in enumerate(self.L.args):
            if i != j and isinstance(b, ArgDat) and b.dat == a.dat and b.acc != "OP_READ":
                return False
        return True

    def _dat_arg(self, pre, post, j, a, inc=None, rows=False):
        """Code for dat arg j around the user kernel's call; returns the
        expression passed for it.  Data goes through the registers b{j}
        (indexed by constants after inlining, so they stay registers): the
        increments of an indirect OP_INC (inc: "atomic" adds them to the dat
        after the call, atomically or not as the launch says; "lds" stores
        them to the block's LDS slots), staged args (_staged), and every arg
        under SoA.  An AoS arg that is not staged is passed as a pointer into
        the dat.  LAY is the kernel's template parameter (0: every
        multi-component dat AoS, 1: all SoA, 2: mixed, each arg by its own
        stride and always through registers), so the layout choice folds at
        compile time."""
        T, D = a.typ, a.dim
        # rows: the owner kernel reads instance-ordered copies of the map rows
        m = self.rep.get(j, j)
        row = f"op2gen_row(A.r{m}, A.h{m}, hb{m}, (size_t)(e0 + j) * A.md{m} + {a.idx})" if rows else \
            f"A.m{m}[(size_t)n * A.md{m} + {a.idx}]"
        tgt = "n" if a.direct else row
        soa = f"(LAY == 1 || (LAY == 2 && A.s{j}))"
        idx = f"({soa} ? t{j} + (size_t)k * A.s{j} : (size_t)t{j} * {D} + k)"
        opt = a.opt is not None
        guard = f"if (A.opt{j}) " if opt else ""
        pre.append(f"    const int t{j} = {f'A.opt{j} ? {tgt} : 0' if opt else tgt};")
        pre.append(f"    {T} b{j}[{D}];")
        writes = a.acc != "OP_READ"
        if inc is not None:
            pre.append(f"    for (int k = 0; k < {D}; ++k) b{j}[k] = ({T})0;")
            ptr = f"b{j}"
            if inc == "atomic":
                post.append(f"    {guard}{{")
                post.append(f"      if (atomic) {{")
                post.append(f"        for (int k = 0; k < {D}; ++k) op2gen_atomic_add(A.d{j} + {idx}, b{j}[k]);")
                post.append(f"      }} else {{")
                post.append(f"        for (int k = 0; k < {D}; ++k) A.d{j}[{idx}] += b{j}[k];")
                post.append(f"      }}")
                post.append(f"    }}")
            else:
                # the chunk's slots, component-major: slot threadIdx.x * ninc + which
                post.append(f"    for (int k = 0; k < {D}; ++k) lds[(size_t)k * lds_s + threadIdx.x * {inc[1]} + {inc[2]}] = b{j}[k];")
        elif self._staged(j, a):
            pre.append(f"    {guard}for (int k = 0; k < {D}; ++k) b{j}[k] = A.d{j}[{idx}];")
            ptr = f"b{j}"
            if writes:
                post.append(f"    {guard}for (int k = 0; k < {D}; ++k) A.d{j}[{idx}] = b{j}[k];")
        else:
            # loaded through registers even for OP_WRITE, so that components
            # the kernel leaves alone are stored back unchanged
            pre.append(f"    if (LAY != 0) {{")
            pre.append(f"      {guard}for (int k = 0; k < {D}; ++k) b{j}[k] = A.d{j}[{idx}];")
            pre.append(f"    }}")
            ptr = f"(LAY != 0 ? b{j} : A.d{j} + (size_t)t{j} * {D})"
            if writes:
                post.append(f"    if (LAY != 0) {{")
                post.append(f"      {guard}for (int k = 0; k < {D}; ++k) A.d{j}[{idx}] = b{j}[k];")
                post.append(f"    }}")
        return f"(A.opt{j} ? {ptr} : nullptr)" if opt else ptr

    # ---- device kernel ----
    def kernel(self):
        L = self.L
        pre, call, post = [], [], []
        for j, a in enumerate(L.args):
            if isinstance(a, ArgDat):
                inc_ind = (not a.direct) and a.acc == "OP_INC"
                call.append(self._dat_arg(pre, post, j, a, "atomic" if inc_ind else None))
            elif isinstance(a, ArgGbl):
                if a.acc == "OP_READ":
                    pre.append(f"    {a.typ} g{j}[{a.dim}];")
                    pre.append(f"    for (int k = 0; k < {a.dim}; ++k) g{j}[k] = A.g{j}[k];")
                    call.append(f"(A.opt{j} ? g{j} : nullptr)" if a.opt is not None else f"g{j}")
                else:
                    call.append(f"r{j}")
            elif isinstance(a, ArgIdx):
                if a.map:
                    m = self.rep.get(j, j)
                    pre.append(f"    int x{j} = A.m{m}[(size_t)n * A.md{m} + {a.idx}];")
                else:
                    pre.append(f"    int x{j} = n;")
                pre.append(f"    if (A.oi{j}) x{j} = A.oi{j}[x{j}];")
                call.append(f"&x{j}")
            elif isinstance(a, ArgInfo):
                call.append(f"v{j}")
        red_decl = []
        for j in self.reds:
            a = L.args[j]
            ident = {"OP_INC": f"({a.typ})0", "OP_MIN": f"op2gen_lim<{a.typ}>::max()",
                     "OP_MAX": f"op2gen_lim<{a.typ}>::lowest()"}[a.acc]
            red_decl.append(f"  {a.typ} r{j}[{a.dim}];")
            red_decl.append(f"  for (int k = 0; k < {a.dim}; ++k) r{j}[k] = {ident};")
            if j in self.info_of:
                ia = L.args[self.info_of[j]]
                red_decl.append(f"  {ia.typ} v{self.info_of[j]}[{ia.dim}] = {{}};")
                red_decl.append(f"  int at{j} = 0x7fffffff;")
        track = []
        for j in self.reds:
            if j in self.info_of:
                track.append(f"    const {L.args[j].typ} before{j} = r{j}[0];")
        # the element a MIN/MAX result came from, for ties: the application's
        # index under transparent renumbering, so the first element in its
        # numbering wins as in the sequential executor
        track_after = [f"    if (r{j}[0] != before{j}) at{j} = A.on ? A.on[n] : n;"
                       for j in self.reds if j in self.info_of]
        body = []
        body.append(f"template <int LAY>")
        body.append(f"__global__ void __launch_bounds__(1024)")
        body.append(f"    {self.name}_kernel({self.name}_args A, const int *__restrict elems, int start,")
        body.append(f"                 int end, int atomic, unsigned char *partials, int part_base,")
        body.append(f"                 int step) {{")
        body.append(f"  const int blk = op2gen_xcd_block();  // XCD-contiguous element ranges")
        body += red_decl
        body.append(f"  // step > 0: the thread's elements are e, e + step, ... (a grid of fewer")
        body.append(f"  // threads than elements, so fewer partial records); 0: one element")
        body.append(f"  for (int e = start + blk * (int)blockDim.x + (int)threadIdx.x; e < end; e += step) {{")
        body.append(f"    const int n = elems ? elems[e] : e;")
        body += pre
        body += track
        body.append(f"    {self.ns}::{L.kernel}({', '.join(call)});")
        body += track_after
        body += post
        body.append(f"    if (step == 0) break;")
        body.append(f"  }}")
        if self.reds:
            # block epilogue: each wave reduces by shuffles, one barrier, then
            # thread 0 combines the waves in wave order (the order of
            # op2gen_block_reduce) and writes the block's partial record
            body.append(f"  {{")
            body.append(f"    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;")
            fin = []
            for j in self.reds:
                a = L.args[j]
                op = {"OP_INC": "0", "OP_MIN": "1", "OP_MAX": "2"}[a.acc]
                o = self.rec_off[j]
                if j in self.info_of:
                    ij = self.info_of[j]
                    ia = L.args[ij]
                    io = o + _pad8(_size(a.typ))
                    ao = io + _pad8(ia.dim * _size(ia.typ))
                    body.append(f"    __shared__ {a.typ} s_v{j}[16];")
                    body.append(f"    __shared__ int s_a{j}[16];")
                    body.append(f"    __shared__ {ia.typ} s_i{j}[16][{ia.dim}];")
                    body.append(f"    {{")
                    body.append(f"      {a.typ} bv = r{j}[0];")
                    body.append(f"      int ba = at{j};")
                    body.append(f"      const int win = op2gen_wave_argbest<{a.typ}, {op}>(bv, ba);")
                    body.append(f"      if (lane == win) {{")
                    body.append(f"        s_v{j}[w] = bv;")
                    body.append(f"        s_a{j}[w] = ba;")
                    body.append(f"        for (int k = 0; k < {ia.dim}; ++k) s_i{j}[w][k] = v{ij}[k];")
                    body.append(f"      }}")
                    body.append(f"    }}")
                    fin.append(f"      {{")
                    fin.append(f"        int bw = 0;")
                    fin.append(f"        for (int i = 1; i < nw; ++i)")
                    fin.append(f"          if (op2gen_better<{a.typ}, {op}>(s_v{j}[i], s_a{j}[i], s_v{j}[bw], s_a{j}[bw])) bw = i;")
                    fin.append(f"        *({a.typ} *)(rec + {o}) = s_v{j}[bw];")
                    fin.append(f"        for (int k = 0; k < {ia.dim}; ++k) (({ia.typ} *)(rec + {io}))[k] = s_i{j}[bw][k];")
                    fin.append(f"        *(int *)(rec + {ao}) = s_a{j}[bw];")
                    fin.append(f"      }}")
                else:
                    body.append(f"    __shared__ {a.typ} s_r{j}[16][{a.dim}];")
                    body.append(f"    for (int k = 0; k < {a.dim}; ++k) {{")
                    body.append(f"      const {a.typ} v = op2gen_wave_reduce<{a.typ}, {op}>(r{j}[k]);")
                    body.append(f"      if (lane == 0) s_r{j}[w][k] = v;")
                    body.append(f"    }}")
                    fin.append(f"      for (int k = 0; k < {a.dim}; ++k) {{")
                    fin.append(f"        {a.typ} acc = s_r{j}[0][k];")
                    fin.append(f"        for (int i = 1; i < nw; ++i) acc = op2gen_combine<{a.typ}, {op}>(acc, s_r{j}[i][k]);")
                    fin.append(f"        (({a.typ} *)(rec + {o}))[k] = acc;")
                    fin.append(f"      }}")
            body.append(f"    __syncthreads();")
            body.append(f"    if (threadIdx.x == 0) {{")
            body.append(f"      unsigned char *rec = partials + (size_t)(part_base + blk) * {self.rec};")
            body += fin
            body.append(f"    }}")
            body.append(f"  }}")
        body.append("}")
        return "\n".join(body) + "\n"

    def owner_kernel(self):
        """Owner-partition variant.  A block owns at most blockDim targets (the
        plan's patches hold at most that many) and walks its instances in
        chunks of blockDim: each thread evaluates one instance of the chunk
        with its increments in registers and stores them in the chunk's LDS
        slots (component-major, so a wave's stores are consecutive); then each
        owned target adds the entries of its list that fall in the chunk.  The
        lists are ascending and chunks are taken in order, so every target
        sums its increments in element order, as the sequential executor
        does.  LDS is blockDim * ninc * dim values whatever the patch's
        perimeter.  Blocks take patches in XCD-contiguous order."""
        if not self.owner:
            return ""
        L = self.L
        incs = [j for j, a in enumerate(L.args) if isinstance(a, ArgDat) and not a.direct
                and a.acc == "OP_INC"]
        T, D, ninc = L.args[incs[0]].typ, L.args[incs[0]].dim, len(incs)
        pre, call, post = [], [], []
        for j, a in enumerate(L.args):
            if isinstance(a, ArgDat):
                inc = ("lds", ninc, incs.index(j)) if j in incs else None
                call.append(self._dat_arg(pre, post, j, a, inc, rows=True))
            elif isinstance(a, ArgGbl):
                pre.append(f"    {a.typ} g{j}[{a.dim}];")
                pre.append(f"    for (int k = 0; k < {a.dim}; ++k) g{j}[k] = A.g{j}[k];")
                call.append(f"(A.opt{j} ? g{j} : nullptr)" if a.opt is not None else f"g{j}")
            elif isinstance(a, ArgIdx):
                if a.map:
                    m = self.rep.get(j, j)
                    pre.append(f"    int x{j} = op2gen_row(A.r{m}, A.h{m}, hb{m}, (size_t)(e0 + j) * A.md{m} + {a.idx});")
                else:
                    pre.append(f"    int x{j} = n;")
                pre.append(f"    if (A.oi{j}) x{j} = A.oi{j}[x{j}];")
                call.append(f"&x{j}")
        w = incs[0]
        at = f"(LAY == 1 || (LAY == 2 && A.s{w})) ? t + (size_t)k * A.s{w} : (size_t)t * {D} + k"
        b = [f"template <int LAY>",
             f"__global__ void __launch_bounds__(1024)",
             f"    {self.name}_owner({self.name}_args A, const int *__restrict elem_offs,",
             f"                 const int *__restrict elems, const int *__restrict owned_offs,",
             f"                 const int *__restrict owned, const int *__restrict csr_offs,",
             f"                 const unsigned short *__restrict csr,",
             f"                 const ushort4 *__restrict csr4) {{",
             f"  extern __shared__ unsigned char lds_raw[];",
             f"  {T} *lds = reinterpret_cast<{T} *>(lds_raw);",
             f"  const int lds_s = (int)blockDim.x * {ninc};  // slots per chunk",
             f"  const int blk = op2gen_xcd_block();",
             f"  const int e0 = elem_offs[blk], ne = elem_offs[blk + 1] - e0;",
             f"  const int o0 = owned_offs[blk], no = owned_offs[blk + 1] - o0;"]
        for j, a in enumerate(L.args):
            if (isinstance(a, ArgDat) and not a.direct) or (isinstance(a, ArgIdx) and a.map):
                b.append(f"  const int hb{j} = A.hb{j} ? A.hb{j}[blk] : -1;  // -1: int rows")
        b += [
             f"  const bool own = (int)threadIdx.x < no;",
             f"  const int t = own ? owned[o0 + threadIdx.x] : 0;",
             f"  {T} acc[{D}];",
             f"  unsigned short es[4] = {{0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF}};",
             f"  int c = 0, c1 = 0;",
             f"  if (own) {{",
             f"    for (int k = 0; k < {D}; ++k) acc[k] = A.d{w}[{at}];",
             f"    if (csr4) {{  // fixed-width lists: one 8-byte load",
             f"      const ushort4 e = csr4[o0 + threadIdx.x];",
             f"      es[0] = e.x; es[1] = e.y; es[2] = e.z; es[3] = e.w;",
             f"    }} else {{",
             f"      c = csr_offs[o0 + threadIdx.x];",
             f"      c1 = csr_offs[o0 + threadIdx.x + 1];",
             f"    }}",
             f"  }}",
             f"  for (int base = 0; base < ne; base += blockDim.x) {{",
             f"    const int j = base + (int)threadIdx.x;",
             f"    if (j < ne) {{",
             f"      const int n = elems[e0 + j];"]
        b += ["  " + ln for ln in pre]
        b.append(f"      {self.ns}::{L.kernel}({', '.join(call)});")
        b += ["  " + ln for ln in post]
        b += [f"    }}",
              f"    __syncthreads();",
              f"    if (own) {{",
              f"      const int lo = base * {ninc}, hi = lo + lds_s;",
              f"      if (csr4) {{",
              f"        for (int q = 0; q < 4; ++q)",
              f"          if (es[q] != 0xFFFF && es[q] >= lo && es[q] < hi)",
              f"            for (int k = 0; k < {D}; ++k) acc[k] += lds[(size_t)k * lds_s + (es[q] - lo)];",
              f"      }} else {{",
              f"        for (; c < c1 && (int)csr[c] < hi; ++c)",
              f"          for (int k = 0; k < {D}; ++k) acc[k] += lds[(size_t)k * lds_s + ((int)csr[c] - lo)];",
              f"      }}",
              f"    }}",
              f"    __syncthreads();",
              f"  }}",
              f"  if (own)",
              f"    for (int k = 0; k < {D}; ++k) A.d{w}[{at}] = acc[k];",
              f"}}"]
        return "\n".join(b) + "\n"

    def finish(self):
        """One workgroup combines the per-block partial records: a single pass
        over the records gathers every reduction's per-thread value (records
        in ascending order per thread), then one workgroup reduction each."""
        if not self.reds:
            return ""
        L = self.L
        params = ", ".join(f"unsigned char *slot{j}" for j in self.reds)
        b = [f"__global__ void __launch_bounds__(1024)",
             f"    {self.name}_finish(const unsigned char *__restrict partials, int nparts, {params}) {{"]
        decl, load, fin = [], [], []
        for j in self.reds:
            a = L.args[j]
            op = {"OP_INC": "0", "OP_MIN": "1", "OP_MAX": "2"}[a.acc]
            o = self.rec_off[j]
            ident = {"OP_INC": f"({a.typ})0", "OP_MIN": f"op2gen_lim<{a.typ}>::max()",
                     "OP_MAX": f"op2gen_lim<{a.typ}>::lowest()"}[a.acc]
            if j in self.info_of:
                ia = L.args[self.info_of[j]]
                io = o + _pad8(_size(a.typ))
                ao = io + _pad8(ia.dim * _size(ia.typ))
                decl.append(f"  {a.typ} bv{j} = {ident}; int ba{j} = 0x7fffffff, bp{j} = -1;")
                load.append(f"    {{")
                load.append(f"      const {a.typ} v = *(const {a.typ} *)(rec + {o});")
                load.append(f"      const int at = *(const int *)(rec + {ao});")
                load.append(f"      if (op2gen_better<{a.typ}, {op}>(v, at, bv{j}, ba{j})) {{ bv{j} = v; ba{j} = at; bp{j} = p; }}")
                load.append(f"    }}")
                fin.append(f"  {{")
                fin.append(f"    const int win = op2gen_block_argbest<{a.typ}, {op}>(bv{j}, ba{j});")
                fin.append(f"    if (threadIdx.x == (unsigned)win) {{")
                fin.append(f"      *({a.typ} *)slot{j} = bv{j};")
                fin.append(f"      if (bp{j} >= 0) for (int k = 0; k < {ia.dim}; ++k)")
                fin.append(f"        (({ia.typ} *)(slot{j} + {_size(a.typ)}))[k] = "
                           f"((const {ia.typ} *)(partials + (size_t)bp{j} * {self.rec} + {io}))[k];")
                fin.append(f"    }}")
                fin.append(f"  }}")
            else:
                decl.append(f"  {a.typ} acc{j}[{a.dim}];")
                decl.append(f"  for (int k = 0; k < {a.dim}; ++k) acc{j}[k] = {ident};")
                load.append(f"    for (int k = 0; k < {a.dim}; ++k)")
                load.append(f"      acc{j}[k] = op2gen_combine<{a.typ}, {op}>(acc{j}[k], "
                            f"((const {a.typ} *)(rec + {o}))[k]);")
                fin.append(f"  for (int k = 0; k < {a.dim}; ++k) {{")
                fin.append(f"    const {a.typ} v = op2gen_block_reduce<{a.typ}, {op}>(acc{j}[k]);")
                fin.append(f"    if (threadIdx.x == 0) (({a.typ} *)slot{j})[k] = v;")
                fin.append(f"  }}")
        b += decl
        b.append(f"#pragma unroll 4")
        b.append(f"  for (int p = threadIdx.x; p < nparts; p += blockDim.x) {{")
        b.append(f"    const unsigned char *rec = partials + (size_t)p * {self.rec};")
        b += load
        b.append(f"  }}")
        b += fin
        b.append("}")
        return "\n".join(b) + "\n"

    def stub(self):
        L = self.L
        N = len(L.args)
        sig = ", ".join(["char const *name", "op_set set"] + [f"op_arg arg{j}" for j in range(N)])
        s = [f'extern "C" void {L.stub}({sig}) {{']
        s.append(f"  op_arg args[{max(N, 1)}] = {{{', '.join(f'arg{j}' for j in range(N)) or '{}'}}};")
        s.append(f"  static bool arith_given = false;  // the cost model's arithmetic feature")
        s.append(f"  if (!arith_given) {{")
        s.append(f"    op2h_set_loop_arith(name, {self.arith});")
        s.append(f"    arith_given = true;")
        s.append(f"  }}")
        s.append(f"  op2h_loop_ctx ctx;")
        s.append(f"  op2h_loop_begin_mask(&ctx, {L.index - 1}, name, set, {N}, args, {self.def_plan}, 256, "
                 f"{self.mask}u);")
        s.append(f"  if (set->size > 0) {{")
        s.append(f"    hipStream_t st = (hipStream_t)ctx.stream;")
        s.append(f"    {self.name}_args A;")
        s.append(f"    memset(&A, 0, sizeof(A));")
        s.append(f"    A.on = op2h_set_old_index(set);")
        for j, a in enumerate(L.args):
            if isinstance(a, ArgDat):
                s.append(f"    A.opt{j} = args[{j}].opt;")
                s.append(f"    A.d{j} = ({a.typ} *)args[{j}].data_d;")
                s.append(f"    A.s{j} = args[{j}].opt ? op2h_dat_soa_stride(args[{j}].dat) : 0;")
                if not a.direct:
                    s.append(f"    A.m{j} = args[{j}].opt ? args[{j}].map_data_d : nullptr;")
                    s.append(f"    A.md{j} = args[{j}].opt ? args[{j}].map->dim : 0;")
            elif isinstance(a, ArgGbl) and a.acc == "OP_READ":
                s.append(f"    A.opt{j} = args[{j}].opt;")
                s.append(f"    if (args[{j}].opt)")
                s.append(f"      for (int k = 0; k < {a.dim}; ++k) A.g{j}[k] = (({a.typ} *)args[{j}].data)[k];")
            elif isinstance(a, ArgIdx) and a.map:
                s.append(f"    A.m{j} = args[{j}].map_data_d;")
                s.append(f"    A.md{j} = args[{j}].map->dim;")
            if isinstance(a, ArgIdx):
                tgt = f"args[{j}].map->to" if a.map else "set"
                s.append(f"    A.oi{j} = op2h_set_old_index({tgt});")
        # the kernels are instantiated for all-AoS, all-SoA and mixed layouts
        # of the loop's multi-component dats
        s.append(f"    int layout = -1;  // 0 AoS, 1 SoA, 2 mixed, -1 no multi-component dat")
        for j, a in enumerate(L.args):
            if isinstance(a, ArgDat) and a.dim > 1:
                s.append(f"    if (args[{j}].opt) {{")
                s.append(f"      const int l = A.s{j} != 0;")
                s.append(f"      layout = layout < 0 || layout == l ? l : 2;")
                s.append(f"    }}")

        s.append(f"    const int bs = ctx.block_size;")
        if self.owner:
            L2 = self.L
            incs = [j for j, a in enumerate(L2.args) if isinstance(a, ArgDat) and not a.direct
                    and a.acc == "OP_INC"]
            T, D = L2.args[incs[0]].typ, L2.args[incs[0]].dim
            s.append(f"    if (ctx.plan == {PLAN_OWNER}) {{")
            s.append(f"      ctx.part_size = bs;")
            s.append(f"      const op2h_owner_plan *op = op2h_plan_owner(&ctx);")
            s.append(f"      if (!op) exit(-1);")
            s.append(f"      const size_t lds = (size_t)bs * {len(incs)} * {D} * sizeof({T});")
            s.append(f"      if (lds > 160 * 1024 || op->max_owned > bs) {{")
            s.append(f'        printf("op2hip: %s: owner plan with %d-target blocks needs %zu bytes of LDS; use a smaller block\\n",')
            s.append(f"               name, op->max_owned, lds);")
            s.append(f"        exit(-1);")
            s.append(f"      }}")
            # instance-ordered copies of each map's rows, packed once per plan
            by_map = {}
            for j, a in enumerate(L2.args):
                if (isinstance(a, ArgDat) and not a.direct) or (isinstance(a, ArgIdx) and a.map):
                    by_map.setdefault(a.map, []).append(j)
            for mp, js in by_map.items():
                s.append(f"      {{")
                s.append(f"        const op_arg *ra = nullptr;  // the first enabled arg on this map")
                for j in js:
                    s.append(f"        if (!ra && args[{j}].opt) ra = &args[{j}];")
                s.append(f"        if (ra) {{")
                s.append(f"          char key[160];")
                s.append(f'          snprintf(key, sizeof key, "owner_rows.%p.%d", (const void *)op, ra->map->index);')
                s.append(f"          int fresh = 0;")
                s.append(f"          const int mdim = ra->map->dim;")
                s.append(f"          const size_t nr = (size_t)op->total * mdim;")
                s.append(f"          // [int rows][int base per block][ushort rows][pad]")
                s.append(f"          unsigned char *mem = (unsigned char *)op2h_plan_user(")
                s.append(f"              key, 6 * nr + 4 * (size_t)op->nblocks + 16, &fresh);")
                s.append(f"          int *rows = (int *)mem, *base = rows + nr;")
                s.append(f"          unsigned short *rows16 = (unsigned short *)(base + op->nblocks);")
                s.append(f"          if (fresh && op->total) {{")
                s.append(f"            op2gen_launch(false, nullptr, nullptr, op2gen_instance_rows, (op->total + 255) / 256, 256,")
                s.append(f"                          0, st, op->elems, op->total, ra->map_data_d, mdim, rows);")
                s.append(f"            op2gen_launch(false, nullptr, nullptr, op2gen_rows16, op->nblocks, 256, 0, st,")
                s.append(f"                          op->elem_offs, rows, mdim, base, rows16);")
                s.append(f"          }}")
                for j in js:
                    s.append(f"          A.r{j} = rows;")
                    s.append(f"          A.h{j} = rows16;")
                    s.append(f"          A.hb{j} = op->total ? base : nullptr;")
                s.append(f"        }}")
                s.append(f"      }}")
            s.append(f"      void *t0 = nullptr, *t1 = nullptr;")
            s.append(f"      const bool ext = op->nblocks && op2h_kernel_timer_events(&ctx, &t0, &t1) != 0;")
            s.append(f"      if (!ext) op2h_kernel_timer_start(&ctx);")
            s.append(f"      if (op->nblocks)")
            s.append(f"        op2gen_launch(ext, t0, t1, layout == 2 ? {self.name}_owner<2> : layout == 1 ?")
            s.append(f"                      {self.name}_owner<1> : {self.name}_owner<0>,")
            s.append(f"                      op->nblocks, bs, lds, st, A,")
            s.append(f"                      op->elem_offs, op->elems, op->owned_offs, op->owned,")
            s.append(f"                      op->csr_offs, op->csr, (const ushort4 *)op->csr4);")
            s.append(f"      if (!ext) op2h_kernel_timer_stop(&ctx);")
            s.append(f"      OP2GEN_CHECK(hipGetLastError());")
            s.append(f"      OP_kernels[{L2.index - 1}].transfer += (float)op2h_algorithmic_bytes(set, {len(L2.args)}, args);")
            s.append(f"      op2h_loop_end(&ctx);")
            s.append(f"      return;")
            s.append(f"    }}")
        s.append(f"    const op2h_color_plan *cp = ctx.plan == {PLAN_COLOR} ? op2h_plan_color(&ctx) : nullptr;")
        s.append(f"    int nparts = 0;")
        s.append(f"    if (cp) {{")
        s.append(f"      for (int c = 0; c < cp->ncolors; ++c)")
        s.append(f"        nparts += (cp->col_offsets_h[c + 1] - cp->col_offsets_h[c] + bs - 1) / bs;")
        s.append(f"    }} else {{")
        s.append(f"      nparts = (set->size + bs - 1) / bs;")
        if self.reds:
            # about three elements per thread, at most 2048 workgroups: fewer
            # partial records for the finish to combine
            s.append(f"      nparts = std::min((set->size + 3 * bs - 1) / (3 * bs), 2048);")
        s.append(f"    }}")
        s.append(f"    const int step = !cp && nparts * bs < set->size ? nparts * bs : 0;")
        if self.reds:
            s.append(f"    unsigned char *partials = (unsigned char *)op2h_reduct_scratch((size_t)nparts * {self.rec});")
        else:
            s.append(f"    unsigned char *partials = nullptr;")
        s.append(f"    auto kern = layout == 2 ? {self.name}_kernel<2> : layout == 1 ? {self.name}_kernel<1> :")
        s.append(f"                {self.name}_kernel<0>;")
        s.append(f"    // one launch (two with a reduction's finish) outside colouring: timed by")
        s.append(f"    // events attached to the kernels; colour launches by an event pair")
        s.append(f"    void *t0 = nullptr, *t1 = nullptr;")
        s.append(f"    const bool ext = !cp && nparts > 0 && op2h_kernel_timer_events(&ctx, &t0, &t1) != 0;")
        s.append(f"    if (!ext) op2h_kernel_timer_start(&ctx);")
        s.append(f"    if (cp) {{")
        s.append(f"      int base = 0;")
        s.append(f"      for (int c = 0; c < cp->ncolors; ++c) {{")
        s.append(f"        const int s0 = cp->col_offsets_h[c], s1 = cp->col_offsets_h[c + 1];")
        s.append(f"        const int g = (s1 - s0 + bs - 1) / bs;")
        s.append(f"        if (g) op2gen_launch(false, nullptr, nullptr, kern, g, bs, 0, st, A, cp->col_reord,")
        s.append(f"                             s0, s1, 0, partials, base, 0);")
        s.append(f"        base += g;")
        s.append(f"      }}")
        s.append(f"    }} else if (nparts > 0) {{")
        stop = "nullptr" if self.reds else "t1"
        s.append(f"      op2gen_launch(ext, t0, {stop}, kern, nparts, bs, 0, st, A,")
        s.append(f"                    nullptr, 0, set->size, 1, partials, 0, step);")
        s.append(f"    }}")
        if self.reds:
            slots = ", ".join(f"(unsigned char *)op2h_gbl_slot(&ctx, {j}, {self.info_of.get(j, -1)})"
                              for j in self.reds)
            # one workgroup of 256: at most 2048 records (8 per thread) outside
            # colouring; fewer waves make its barriers and wave combines cheaper
            s.append(f"    op2gen_launch(ext, nullptr, t1, {self.name}_finish, 1, 256, 0, st, partials,")
            s.append(f"                  nparts, {slots});")
        s.append(f"    if (!ext) op2h_kernel_timer_stop(&ctx);")
        s.append(f"    OP2GEN_CHECK(hipGetLastError());")
        s.append(f"  }}")
        s.append(f"  OP_kernels[{L.index - 1}].transfer += (float)op2h_algorithmic_bytes(set, {N}, args);")
        s.append(f"  op2h_loop_end(&ctx);")
        s.append(f"}}")
        return "\n".join(s) + "\n"


PRELUDE = r'''#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <tuple>
#include <utility>

#include "op2_hip.h"
#include "op2_hip_rt.h"

#define OP2GEN_CHECK(x)                                                      \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__,     \
             hipGetErrorString(e_));                                         \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

namespace {

// Launch with the kernel's declared argument types; with timing events
// (op2h_kernel_timer_events) through hipExtLaunchKernel, which stamps e0 at
// the kernel's start and e1 at its end (either may be NULL).
template <class Tuple, std::size_t... I>
void op2gen_arg_ptrs(Tuple &t, void **p, std::index_sequence<I...>) {
  ((p[I] = (void *)&std::get<I>(t)), ...);
}

template <typename... KArgs, typename... Args>
void op2gen_launch(bool ext, void *e0, void *e1, void (*kernel)(KArgs...), dim3 grid,
                   dim3 block, size_t lds, hipStream_t st, Args... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  std::tuple<KArgs...> t{static_cast<KArgs>(args)...};
  void *p[sizeof...(KArgs) > 0 ? sizeof...(KArgs) : 1];
  op2gen_arg_ptrs(t, p, std::index_sequence_for<KArgs...>{});
  if (ext)
    OP2GEN_CHECK(hipExtLaunchKernel((const void *)kernel, grid, block, p, lds, st,
                                    (hipEvent_t)e0, (hipEvent_t)e1, 0));
  else
    OP2GEN_CHECK(hipLaunchKernel((const void *)kernel, grid, block, p, lds, st));
}

// Instance j's row of a map (the row of element elems[j]), packed once per
// owner plan so that the owner kernel reads one row per instance directly.
__global__ void op2gen_instance_rows(const int *__restrict elems, int total,
                                     const int *__restrict map, int dim,
                                     int *__restrict rows) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total) return;
  const int e = elems[j];
  for (int d = 0; d < dim; ++d) rows[(size_t)j * dim + d] = map[(size_t)e * dim + d];
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8), each
// with its own L2.  Logical block ids give each XCD one contiguous range, so
// neighbouring patches of a locally numbered mesh share an L2.  A bijection of
// [0, gridDim.x).
__device__ __forceinline__ int op2gen_xcd_block() {
  constexpr int nxcd = 8;
  const int b = blockIdx.x, n = gridDim.x;
  if (n < nxcd) return b;
  const int q = n / nxcd, r = n % nxcd;
  const int x = b % nxcd, i = b / nxcd;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Per block b of an owner plan (one workgroup of 256 threads each): base[b]
// = the least entry of the block's instance rows when every entry lies within
// 65535 of it, rows16 then holding entry - base, which halves the bytes an
// owner kernel reads per row; else base[b] = -1 and the block reads the int
// rows.
__global__ void __launch_bounds__(256)
    op2gen_rows16(const int *__restrict elem_offs, const int *__restrict rows, int dim,
                  int *__restrict base, unsigned short *__restrict rows16) {
  __shared__ int s_mn[4], s_mx[4];
  const int b = blockIdx.x;
  const size_t lo = (size_t)elem_offs[b] * dim, hi = (size_t)elem_offs[b + 1] * dim;
  int mn = 0x7fffffff, mx = -1;
  for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    mn = min(mn, rows[i]);
    mx = max(mx, rows[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn = min(mn, __shfl_xor(mn, o, 64));
    mx = max(mx, __shfl_xor(mx, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    s_mn[threadIdx.x >> 6] = mn;
    s_mx[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  mn = min(min(s_mn[0], s_mn[1]), min(s_mn[2], s_mn[3]));
  mx = max(max(s_mx[0], s_mx[1]), max(s_mx[2], s_mx[3]));
  const bool narrow = hi == lo || (long long)mx - mn < 65535;
  if (threadIdx.x == 0) base[b] = hi == lo ? 0 : narrow ? mn : -1;
  if (narrow)
    for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
      rows16[i] = (unsigned short)(rows[i] - mn);
}

// Entry i of an owner kernel's instance rows: from the 16-bit rows when the
// block has a base (uniform over the workgroup), else from the int rows.  One
// 32-bit load either way (the aligned word holding h[i]; the 16-bit rows
// start 4-byte aligned and are padded), its address selected without a
// branch, so the loads of all rows of an instance issue together.
__device__ __forceinline__ int op2gen_row(const int *__restrict r,
                                          const unsigned short *__restrict h, int base,
                                          size_t i) {
  const bool narrow = base >= 0;
  const unsigned *w = narrow ? reinterpret_cast<const unsigned *>(h + (i & ~(size_t)1))
                             : reinterpret_cast<const unsigned *>(r + i);
  const unsigned v = *w;
  return narrow ? base + (int)((i & 1) ? v >> 16 : v & 0xFFFFu) : (int)v;
}

template <class T> struct op2gen_lim {
  __device__ static T max() { return std::numeric_limits<T>::max(); }
  __device__ static T lowest() { return std::numeric_limits<T>::lowest(); }
};

__device__ __forceinline__ void op2gen_atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void op2gen_atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void op2gen_atomic_add(int *p, int v) { atomicAdd(p, v); }

// OP: 0 sum, 1 min, 2 max
template <class T, int OP> __device__ __forceinline__ T op2gen_combine(T a, T b) {
  return OP == 0 ? a + b : OP == 1 ? (b < a ? b : a) : (b > a ? b : a);
}

// (value, element) pairs: the better value wins, ties go to the lower element
template <class T, int OP>
__device__ __forceinline__ bool op2gen_better(T v, int at, T bv, int ba) {
  return OP == 1 ? (v < bv || (v == bv && at < ba)) : (v > bv || (v == bv && at < ba));
}

// Workgroup reduction: xor-shuffle tree inside each wave, then the waves'
// results in wave order (a fixed order for a launch shape).  Lanes past the
// end of a partial last wave take no part.
template <class T, int OP> __device__ T op2gen_block_reduce(T v) {
  __shared__ T part[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = (blockDim.x + 63) >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    const T u = __shfl_xor(v, o);
    if ((t ^ o) < (int)blockDim.x) v = op2gen_combine<T, OP>(v, u);
  }
  if (lane == 0) part[w] = v;
  __syncthreads();
  T r = part[0];
  for (int i = 1; i < nw; ++i) r = op2gen_combine<T, OP>(r, part[i]);
  __syncthreads();
  return r;
}

// Xor-shuffle tree inside a wave (the first stage of op2gen_block_reduce);
// lanes past the end of a partial last wave take no part.
template <class T, int OP> __device__ __forceinline__ T op2gen_wave_reduce(T v) {
  const int t = threadIdx.x;
  if ((blockDim.x & 63) == 0) {  // whole waves (uniform): no lane guard
    for (int o = 32; o > 0; o >>= 1) v = op2gen_combine<T, OP>(v, __shfl_xor(v, o));
    return v;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const T u = __shfl_xor(v, o);
    if ((t ^ o) < (int)blockDim.x) v = op2gen_combine<T, OP>(v, u);
  }
  return v;
}

// Wave arg-best over (value, element) pairs: v and at become the wave's best
// pair; returns the lane that held it, as lane 0 sees it (lanes tied on an
// equal pair may each name themselves), so exactly one lane is named.
template <class T, int OP> __device__ __forceinline__ int op2gen_wave_argbest(T &v, int &at) {
  const int t = threadIdx.x;
  const bool whole = (blockDim.x & 63) == 0;  // uniform
  int who = t & 63;
  for (int o = 32; o > 0; o >>= 1) {
    const T ov = __shfl_xor(v, o);
    const int oa = __shfl_xor(at, o), ow = __shfl_xor(who, o);
    const bool take = (whole || (t ^ o) < (int)blockDim.x) && op2gen_better<T, OP>(ov, oa, v, at);
    v = take ? ov : v;
    at = take ? oa : at;
    who = take ? ow : who;
  }
  return __shfl(who, 0);
}

// Workgroup arg-best over (value, element) pairs: returns the winning thread.
template <class T, int OP> __device__ int op2gen_block_argbest(T v, int at) {
  __shared__ T bv[16];
  __shared__ int ba[16], bt[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = (blockDim.x + 63) >> 6;
  int who = t;
  for (int o = 32; o > 0; o >>= 1) {
    const T ov = __shfl_xor(v, o);
    const int oa = __shfl_xor(at, o), ow = __shfl_xor(who, o);
    if ((t ^ o) < (int)blockDim.x && op2gen_better<T, OP>(ov, oa, v, at)) {
      v = ov;
      at = oa;
      who = ow;
    }
  }
  if (lane == 0) {
    bv[w] = v;
    ba[w] = at;
    bt[w] = who;
  }
  __syncthreads();
  T rv = bv[0];
  int ra = ba[0], rw = bt[0];
  for (int i = 1; i < nw; ++i)
    if (op2gen_better<T, OP>(bv[i], ba[i], rv, ra)) {
      rv = bv[i];
      ra = ba[i];
      rw = bt[i];
    }
  __syncthreads();
  return rw;
}

}  // namespace
'''


def generate(prog):
    """(kernel file text, header text) for a parsed program."""
    ns = f"op2gen_{prog.stem}"
    out = [f"// Generated by op2amd.translator from {prog.stem}: {len(prog.loops)} loops.\n",
           PRELUDE, f"namespace {ns} {{\n"]
    for c in prog.consts:
        out.append(f"__constant__ {c.typ} {c.ptr}" + (f"[{c.dim}]" if c.dim > 1 else "") + ";\n")
    out.append("\n")
    for name, text in prog.device_text.items():
        out.append(_device_kernel(text) + "\n\n")
    out.append(f"}}  // namespace {ns}\n\n")
    # constants hook
    out.append(f"static int {ns}_consts(int dim, const char *type, int size, const char *dat,\n"
               f"                        const char *name) {{\n  (void)type;\n")
    for c in prog.consts:
        out.append(f'  if (!strcmp(name, "{c.ptr}")) {{\n'
                   f"    OP2GEN_CHECK(hipMemcpyToSymbol(HIP_SYMBOL({ns}::{c.ptr}), dat, (size_t)dim * size));\n"
                   f"    return 1;\n  }}\n")
    out.append("  return 0;\n}\n\n")
    out.append(f"static const struct {ns}_hook {{\n  {ns}_hook() {{ op2h_register_const_hook({ns}_consts); }}\n"
               f"}} {ns}_hook_v;\n\n")
    for loop in prog.loops:
        g = LoopGen(loop, ns, _arith_estimate(prog.kernel_text[loop.kernel]))
        out.append(f"// ---- loop {loop.index}: {loop.kernel} ----\n")
        out.append("namespace {\n")
        out.append(g.args_struct())
        out.append(g.kernel())
        out.append(g.owner_kernel())
        out.append(g.finish())
        out.append("}  // namespace\n\n")
        out.append(g.stub())
        out.append("\n")
    guard = f"{ns.upper()}_H"
    hdr = [f"// Generated by op2amd.translator: stubs of {prog.stem}.\n",
           f"#ifndef {guard}\n#define {guard}\n#include \"op2_hip.h\"\n",
           '#ifdef __cplusplus\nextern "C" {\n#endif\n']
    for loop in prog.loops:
        hdr.append(f"void {loop.stub}(char const *, op_set" + ", op_arg" * len(loop.args) + ");\n")
    hdr.append("#ifdef __cplusplus\n}\n#endif\n#endif\n")
    return "".join(out), "".join(hdr)
