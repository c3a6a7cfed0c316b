"""The hand-scheduled main loop of the CRT residue-product kernel (tiled-mm_amd/csrc/tmm_crt_loop.inc,
run by crt_product_kernel in tmm_ozaki.hip), checked on the CPU by executing its instruction
stream for every wave of a workgroup:

  * every global load and every LDS access stays inside its buffer (slices padded to 256 rows and
    64 k, LDS ring of S slots of 32 KB);
  * every LDS fragment read of k-step x reads the slot that k-step x was loaded into, after all
    waves' loads of it have landed (vmcnt) and a barrier has followed;
  * a slot is reloaded only after every earlier read of it has completed (lgkmcnt) and a barrier
    has followed;
  * the MFMAs of k-step x read the fragment set that holds k-step x, after its reads completed;
  * the loop runs exactly one k-step per MFMA block (64 MFMAs), for every k-step count.

The model is the GCN counter semantics: vmcnt(N) / lgkmcnt(N) wait until at most N of the
wave's loads / LDS reads are outstanding, both complete in order; s_barrier joins the waves. The
waves run the same instruction stream, so barrier i of one wave meets barrier i of the others.
"""
import os
import re

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "tiled-mm_amd", "csrc")
STAGE, TILE = 32768, 16384


def loop_text(S):
    src = open(os.path.join(CSRC, "tmm_crt_loop.inc")).read()
    block = src.split(f"#define TMM_CRT_LOOP_S{S} \\\n", 1)[1].split('  ""', 1)[0]
    return [l.strip()[1:].split("\\n")[0] for l in block.strip().splitlines()]


def test_include_is_generated():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(CSRC, "gen_crt_loop.py"), "--check"])
    assert r.returncode == 0, "tmm_crt_loop.inc differs from what gen_crt_loop.py writes"


class Wave:
    """One wave's execution of the loop; global loads and LDS accesses recorded as events."""

    def __init__(self, S, wave, nkt, kpad, lanes=64):
        self.S, self.nkt, self.kpad = S, nkt, kpad
        wm, wn = wave & 1, wave >> 1
        # k-step tiled operands: row `local` of a 16 KB block at byte local*64, lane chunk swizzled
        self.vo = [[((wave * 4 + t) * 16 + (lane >> 2)) * 64 + 16 * ((lane & 3) ^ (((lane >> 2) >> 1) & 3))
                    for lane in range(lanes)] for t in range(4)]
        self.ldsw = wave * 4 * 1024
        lane_off = [(l & 15) * 64 + (((l >> 4) ^ (((l & 15) >> 1) & 3)) << 4) for l in range(lanes)]
        self.ra = [[s * STAGE + wm * 8192 + o for o in lane_off] for s in range(S)]
        self.rb = [[s * STAGE + TILE + wn * 8192 + o for o in lane_off] for s in range(S)]
        self.sreg = {"cnt": nkt, "adv": nkt - 1, "tmp": None, "tmp2": None}
        self.m0 = None
        self.events = []  # (kind, ...) in program order

    def run(self, lines):
        labels = {l[:-1]: i for i, l in enumerate(lines) if l.endswith(":")}
        pc, scc, steps = 0, 0, 0
        while pc < len(lines):
            steps += 1
            assert steps < 10 ** 6, "loop does not terminate"
            ins = lines[pc]
            pc += 1
            if ins.endswith(":"):
                continue
            op, _, rest = ins.partition(" ")
            args = [a.strip() for a in rest.split(",")] if rest else []
            name = lambda a: re.fullmatch(r"%\[(\w+)\]", a).group(1)
            if op == "s_add_u32" and args[0] == "m0":
                assert args[1] == "%[ldsw]"
                self.m0 = self.ldsw + int(args[2])
            elif op == "global_load_lds_dwordx4":
                # an immediate offset moves both the global address and the LDS destination
                t = int(name(args[0])[2:])
                g, _, off = args[1].partition(" offset:")
                off = int(off or 0)
                self.events.append(("load", name(g), [v + off for v in self.vo[t]], self.m0 + off))
            elif op == "s_cmp_lg_u32":
                scc = int(self.sreg[name(args[0])] != int(args[1]))
            elif op == "s_cmp_eq_u32":
                scc = int(self.sreg[name(args[0])] == int(args[1]))
            elif op == "s_cselect_b32":
                self.sreg[name(args[0])] = int(args[1]) if scc else int(args[2])
            elif op == "s_sub_u32":
                a, b = args[1], args[2]
                va = self.sreg[name(a)]
                vb = self.sreg[name(b)] if b.startswith("%") else int(b)
                self.sreg[name(args[0])] = va - vb
                assert self.sreg[name(args[0])] >= 0
            elif op == "v_add_u32":
                t = int(name(args[0])[2:])
                assert name(args[2]) == name(args[0])
                inc = self.sreg[name(args[1])]
                self.vo[t] = [v + inc for v in self.vo[t]]
            elif op == "s_cbranch_scc1":
                if scc:
                    pc = labels[args[0]]
            elif op == "s_branch":
                pc = labels[args[0]]
            elif op == "ds_read_b128":
                dst = name(args[0])
                m = re.fullmatch(r"%\[r([ab])(\d+)\] offset:(\d+)", args[1])
                base = (self.ra if m.group(1) == "a" else self.rb)[int(m.group(2))]
                self.events.append(("read", dst, [b + int(m.group(3)) for b in base]))
            elif op == "v_mfma_i32_16x16x64_i8":
                self.events.append(("mfma", name(args[0]), name(args[1]), name(args[2])))
            elif op == "s_waitcnt":
                vm = re.search(r"vmcnt\((\d+)\)", rest)
                lg = re.search(r"lgkmcnt\((\d+)\)", rest)
                self.events.append(("wait", int(vm.group(1)) if vm else None, int(lg.group(1)) if lg else None))
            elif op == "s_barrier":
                self.events.append(("barrier",))
            elif op == "s_nop":
                pass
            else:
                raise AssertionError(f"instruction not modelled: {ins}")
        assert self.sreg["cnt"] == 0


def check(S, nkt):
    kpad = 64 * nkt
    lines = loop_text(S)
    waves = [Wave(S, w, nkt, kpad) for w in range(4)]
    for w in waves:
        w.run(lines)
    buf = 256 * kpad  # bytes of one operand's 256-row block: nkt k-step blocks of 16 KB
    # per wave: annotate loads with k-step index and completion point (barrier epoch it is visible at)
    info = []
    wi_of = {id(w): i for i, w in enumerate(waves)}
    for w in waves:
        loads, reads, mfmas = [], [], []
        out_v, out_l = [], []  # outstanding load / read indices
        epoch = 0
        done_v, done_l = {}, {}
        nload = 0
        for ev in w.events:
            if ev[0] == "load":
                _, g, offs, m0 = ev
                step = nload // 8
                nload += 1
                for o in offs:
                    assert 0 <= o <= buf - 16, f"global load out of bounds: {o}"
                    assert o // TILE == min(step, nkt - 1), "load outside its k-step block"
                assert 0 <= m0 and m0 + 64 * 16 <= S * STAGE, "LDS DMA out of the ring"
                # lane l lands at LDS byte m0 + 16 l: row p // 64, 16-byte slot (p % 64) // 16, which
                # must hold chunk slot ^ swizzle(row) of the same row of the block
                for lane, o in enumerate(offs):
                    p = m0 % TILE + 16 * lane
                    row, slot_ = p // 64, (p % 64) // 16
                    assert (o % TILE) // 64 == row, "load lands in another LDS row"
                    assert (o % 64) // 16 == slot_ ^ ((row >> 1) & 3), "load lands in the wrong swizzle slot"
                slot = m0 // STAGE
                assert (m0 % STAGE) // TILE == (0 if g == "gA" else 1)
                loads.append({"slot": slot, "step": min(step, nkt - 1), "raw": step, "epoch_issued": epoch})
                out_v.append(len(loads) - 1)
            elif ev[0] == "read":
                _, dst, addrs = ev
                for lane, a in enumerate(addrs):
                    assert 0 <= a and a + 16 <= S * STAGE, "LDS read out of the ring"
                    # fragment lane: row (block * 16 + lane % 16) of the wave's half, k chunk lane // 16
                    row, slot_ = (a % TILE) // 64, (a % 64) // 16
                    assert row % 16 == lane % 16 and slot_ ^ ((row >> 1) & 3) == lane // 16
                    half = (wi_of[id(w)] & 1) if (a % STAGE) < TILE else (wi_of[id(w)] >> 1)
                    assert row // 128 == half, "fragment read from the other wave's half"
                reads.append({"slot": addrs[0] // STAGE, "dst": dst, "epoch": epoch, "done": None})
                out_l.append(len(reads) - 1)
            elif ev[0] == "wait":
                _, vm, lg = ev
                if vm is not None:
                    while len(out_v) > vm:
                        done_v[out_v.pop(0)] = epoch
                if lg is not None:
                    while len(out_l) > lg:
                        reads[out_l.pop(0)]["done"] = epoch
            elif ev[0] == "barrier":
                epoch += 1
            elif ev[0] == "mfma":
                mfmas.append((ev[1], ev[2], ev[3], epoch, len(reads), [r["done"] for r in reads]))
        for i, l in enumerate(loads):
            l["landed_epoch"] = done_v.get(i, 10 ** 9)
        info.append((loads, reads, mfmas))
    # cross-wave rules, with barrier epochs: an event at epoch e happens after barrier e-1 of every wave
    n_steps_read = None
    for wi, (loads, reads, mfmas) in enumerate(info):
        # reads: the r-th group of 16 reads is k-step r (the last group reads a duplicate of the last)
        assert len(reads) % 16 == 0
        n_steps_read = len(reads) // 16
        assert n_steps_read == nkt + 1
        for r_i, r in enumerate(reads):
            step = r_i // 16
            want = min(step, nkt - 1)
            assert r["slot"] == step % S, f"k-step {step} read from slot {r['slot']}"
            for (lw, _, _) in info:
                # the latest load into this slot issued before the read (in barrier order) is k-step `want`
                mine = [l for l in lw if l["slot"] == r["slot"] and l["epoch_issued"] < r["epoch"]]
                last = [l for l in mine if l["raw"] == max(x["raw"] for x in mine)]
                assert last and last[0]["step"] == want, f"slot {r['slot']} holds the wrong k-step"
                # all of that wave's loads of the step landed before a barrier the read follows
                for l in lw:
                    if l["slot"] == r["slot"] and l["raw"] == last[0]["raw"]:
                        assert l["landed_epoch"] < r["epoch"], "read before the loads landed and a barrier"
                # and no later load into the slot can land before the read completes
                later = [l for l in lw if l["slot"] == r["slot"] and l["raw"] > last[0]["raw"]]
                for l in later:
                    assert r["done"] is not None and r["done"] < l["epoch_issued"], \
                        "slot reloaded before a barrier after its reads completed"
        # MFMAs: block b of 64 uses the set holding k-step b, all of whose reads completed before
        assert len(mfmas) == 64 * nkt, f"{len(mfmas)} MFMAs for {nkt} k-steps"
        for q, (acc, srca, srcb, ep, nreads, done) in enumerate(mfmas):
            step = q // 64
            group = range(16 * step, 16 * step + 16)
            dsts = {reads[g]["dst"] for g in group}
            assert srca in dsts and srcb in dsts, f"k-step {step} MFMA uses a fragment of another set"
            assert all(done[g] is not None for g in group), "MFMA before its fragment reads completed"
        accs = [m[0] for m in mfmas[:64]]
        assert len(set(accs)) == 64
    return True


@pytest.mark.parametrize("S", [4])
@pytest.mark.parametrize("nkt", [1, 2, 3, 4, 5, 6, 7, 9, 10, 11, 12, 21])
def test_schedule(S, nkt):
    assert check(S, nkt)
