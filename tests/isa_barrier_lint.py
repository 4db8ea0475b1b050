"""Static check of compiled gfx950 assembly: no workgroup barrier inside a
divergent loop.

On a wave64 machine a loop whose exit the compiler treats as per-lane
(divergent) runs with an exec mask: lanes that leave early wait at the loop
exit while the rest of their wave keeps iterating.  An s_barrier inside such a
loop is then executed by a partial wave, out of step with the lanes waiting
outside, and the workgroup can spin for ever even though every thread reaches
the same __syncthreads() calls in the source.  The compiler lowers a divergent
loop latch to

    s_andn2_b64 exec, exec, <exit mask>
    s_cbranch_execnz <header>        (or s_cbranch_execz <exit> + fall through)

so this check parses each kernel's control-flow graph, finds every such latch,
collects the natural loop it closes, and reports the loops that contain an
s_barrier.  Loops whose trip count is wave-uniform end in a scalar branch
(s_cbranch_scc*/vcc*) and are not reported.
"""
import os
import re
import subprocess
import tempfile

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

_LABEL = re.compile(r"^(\.LBB\d+_\d+):")
_BB_COMMENT = re.compile(r"^; %bb\.(\d+):")
_FUNC = re.compile(r"^(_Z\w+):")


def compile_to_asm(src, flags, out_dir=None):
    """Device-only assembly of one HIP source for the flags given."""
    out_dir = out_dir or tempfile.mkdtemp(prefix="isa_")
    out = os.path.join(out_dir, os.path.basename(src) + ".s")
    subprocess.check_call([HIPCC, "-x", "hip", "--cuda-device-only", "-S", *flags, src, "-o", out],
                          stderr=subprocess.DEVNULL)
    with open(out) as f:
        return f.read()


def _functions(asm):
    """{name: [lines]} of each function body in the file."""
    funcs, cur = {}, None
    for raw in asm.splitlines():
        line = raw.strip()
        m = _FUNC.match(line)
        if m:
            cur = m.group(1)
            funcs[cur] = []
            continue
        if cur is None:
            continue
        if line.startswith(".Lfunc_end"):
            cur = None
            continue
        funcs[cur].append(line)
    return funcs


def _blocks(lines):
    """Split a function body into basic blocks: (name, [instructions])."""
    blocks = []
    name, ins = "entry", []
    anon = 0

    def close():
        if ins or not name.startswith("anon"):
            blocks.append((name, ins))

    for line in lines:
        lab = _LABEL.match(line)
        bb = _BB_COMMENT.match(line)
        if lab or bb:
            close()
            name, ins = (lab.group(1) if lab else "%bb." + bb.group(1)), []
            continue
        if not line or line.startswith("."):
            continue
        if line.startswith(";"):
            continue
        op = line.split()[0]
        ins.append(line)
        if op.startswith("s_branch") or op.startswith("s_cbranch") or op in ("s_endpgm", "s_setpc_b64"):
            close()
            anon += 1
            name, ins = f"anon{anon}", []
    close()
    return blocks


def _cfg(blocks):
    names = [b[0] for b in blocks]
    index = {n: i for i, n in enumerate(names)}
    succ = [[] for _ in blocks]
    for i, (_, ins) in enumerate(blocks):
        last = ins[-1] if ins else ""
        op = last.split()[0] if last else ""
        falls = i + 1 < len(blocks)
        if op == "s_branch":
            succ[i].append(index[last.split()[1]])
        elif op.startswith("s_cbranch"):
            succ[i].append(index[last.split()[1]])
            if falls:
                succ[i].append(i + 1)
        elif op in ("s_endpgm", "s_setpc_b64"):
            pass
        elif falls:
            succ[i].append(i + 1)
    return names, succ


def _natural_loop(succ, header, latch):
    pred = [[] for _ in succ]
    for u, vs in enumerate(succ):
        for v in vs:
            pred[v].append(u)
    loop, stack = {header, latch}, [latch]
    while stack:
        u = stack.pop()
        if u == header:
            continue
        for w in pred[u]:
            if w not in loop:
                loop.add(w)
                stack.append(w)
    return loop


def _dominators(succ):
    """Immediate dominators (Cooper, Harvey and Kennedy's iterative method)."""
    n = len(succ)
    order, seen, stack = [], [False] * n, [(0, iter(succ[0]))]
    seen[0] = True
    while stack:
        u, it = stack[-1]
        nxt = next(it, None)
        if nxt is None:
            stack.pop()
            order.append(u)
        elif not seen[nxt]:
            seen[nxt] = True
            stack.append((nxt, iter(succ[nxt])))
    rpo = order[::-1]
    pos = {u: i for i, u in enumerate(rpo)}
    pred = [[] for _ in range(n)]
    for u in rpo:
        for v in succ[u]:
            pred[v].append(u)
    idom = [None] * n
    idom[0] = 0
    changed = True
    while changed:
        changed = False
        for u in rpo[1:]:
            new = None
            for w in pred[u]:
                if idom[w] is None:
                    continue
                if new is None:
                    new = w
                    continue
                a, b = w, new
                while a != b:
                    while pos[a] > pos[b]:
                        a = idom[a]
                    while pos[b] > pos[a]:
                        b = idom[b]
                new = a
            if idom[u] != new:
                idom[u] = new
                changed = True
    return idom


def _dominates(idom, a, b):
    while True:
        if b == a:
            return True
        if b is None or idom[b] == b:
            return False
        b = idom[b]


def divergent_latches(blocks, succ):
    """(latch, header) pairs of loops closed by an exec-mask latch."""
    idom = _dominators(succ)
    out = []
    for i, (_, ins) in enumerate(blocks):
        if len(ins) < 2:
            continue
        last = ins[-1].split()
        if last[0] not in ("s_cbranch_execnz", "s_cbranch_execz"):
            continue
        # the exec update closing the loop, possibly separated by s_nop
        k = len(ins) - 2
        while k >= 0 and ins[k].split()[0] in ("s_nop",):
            k -= 1
        if k < 0 or not re.match(r"s_andn2_b(32|64)\s+exec,\s*exec,", ins[k]):
            continue
        if last[0] == "s_cbranch_execnz":
            header = succ[i][0]
        else:
            if len(succ[i]) < 2:
                continue
            header = succ[i][1]   # fall-through continues the loop
        if idom[i] is not None and _dominates(idom, header, i):
            out.append((i, header))
    return out


def barriers_in_divergent_loops(asm):
    """[(function, header label, latch label, barrier count)] violations."""
    bad = []
    for fname, lines in _functions(asm).items():
        blocks = _blocks(lines)
        names, succ = _cfg(blocks)
        for latch, header in divergent_latches(blocks, succ):
            loop = _natural_loop(succ, header, latch)
            nbar = sum(1 for b in loop for x in blocks[b][1] if x.startswith("s_barrier"))
            if nbar:
                bad.append((fname, names[header], names[latch], nbar))
    return bad


def count_divergent_loops(asm):
    """Number of exec-mask loops per function (sanity check of the parser)."""
    res = {}
    for fname, lines in _functions(asm).items():
        blocks = _blocks(lines)
        _, succ = _cfg(blocks)
        res[fname] = len(divergent_latches(blocks, succ))
    return res


def count_barriers(asm):
    """s_barrier instructions per function (sanity check of the parser)."""
    return {f: sum(1 for x in lines if x.startswith("s_barrier")) for f, lines in _functions(asm).items()}
