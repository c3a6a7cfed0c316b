"""Writes tmm_crt_loop.inc: the main loop of the CRT residue-product kernel (tmm_ozaki.hip) as
one GCN inline-assembly block, for an S-stage LDS ring (S >= 3).

Run from this directory after changing the schedule: python3 gen_crt_loop.py

Why assembly: with one wave per SIMD nothing else hides a wave's own stalls, and the compiler's
schedule for the equivalent builtins waits for all 16 LDS fragment reads of a k-step before its
first MFMA, then for the barrier, leaving the matrix core idle for a large part of every step.
Here the schedule is fixed:

  * two register sets of fragments (F0, F1: 8 A and 8 B fragments of 16 bytes per lane each);
    k-step `it` multiplies out of F[it % 2] while the LDS reads of k-step it+1 fill the other set;
  * k-step it, first half (32 MFMAs, B fragments 0..3): 4 of the 8 global->LDS loads of k-step
    it+S-1, one per 8 MFMAs (M0 is set one MFMA before its load: the LDS DMA reads M0 and needs
    a wait state after an SALU write of it);
  * half-way: wait for this wave's loads of k-step it+1 (vmcnt), then the workgroup barrier, after
    which every wave's share of k-step it+1 is in LDS;
  * second half (32 MFMAs, B fragments 4..7): the 16 LDS reads of k-step it+1 into the other set,
    one after each of the first 16 MFMAs, so the last 16 MFMAs cover their latency; the other 4
    loads between those last 16 MFMAs.

A slot is reloaded only after every wave has passed a barrier that follows its last read of the
slot: the load of k-step it+S-1 goes into slot (it-1) % S, whose reads (k-step it-1, issued in
k-step it-2's second half) every wave finished at the top of k-step it-1, before that step's
barrier. Loads past the last k-step reload the last one (the address stops advancing), so every
load is in bounds and the vmcnt wait needs no tail case.

The operands are in k-step tiled layout (one contiguous 16 KB block per 256 rows and 64 k), so
load t of a wave reads the contiguous 1 KB at vo<t> = vo<0> + t*1024 of the current block, and
the next k-step is 16 KB further. (The loads carry no immediate offset: an LDS-DMA load adds it
to the LDS destination too, so each load gets its own offset register and its own M0.)

Operands (named, see the asm statement in tmm_ozaki.hip): c<i><j> accumulators (AGPR tuples),
fa/fb/ga/gb<i> fragment sets F0 (fa, fb) and F1 (ga, gb), vo<t> the per-lane byte offset of
load t in the current k-step block (shared by A and B), ra<s>/rb<s> per-lane LDS read addresses of slot
s, ldsw the wave's LDS DMA base, gA/gB the tile bases in global memory, cnt the k-step count, adv
the address advances left, tmp/tmp2 scratch SGPRs.
"""
import os
import sys

STAGE = 32768  # bytes of one ring slot (A 16 KB + B 16 KB)
STAGES = (4,)  # ring depths emitted (5 stages, 160 KB of LDS, measured no faster on MI355X)
TILE = 16384


def load_stage(slot):
    """The 8 loads of one k-step into `slot` (4 of A, 4 of B) as (m0 instruction, load
    instruction) pairs, then the address advance."""
    # no immediate offset on the loads: an LDS-DMA load adds it to the LDS address as well
    pairs = []
    for base, g in ((0, "gA"), (TILE, "gB")):
        for t in range(4):
            pairs.append((f"s_add_u32 m0, %[ldsw], {slot * STAGE + base + t * 1024}",
                          f"global_load_lds_dwordx4 %[vo{t}], %[{g}]"))
    advance = [
        "s_cmp_lg_u32 %[adv], 0",
        f"s_cselect_b32 %[tmp], {TILE}, 0",
        "s_cselect_b32 %[tmp2], 1, 0",
        "s_sub_u32 %[adv], %[adv], %[tmp2]",
    ] + [f"v_add_u32 %[vo{t}], %[tmp], %[vo{t}]" for t in range(4)]
    return pairs, advance


def reads(slot, a, b):
    return ([f"ds_read_b128 %[{a}{i}], %[ra{slot}] offset:{i * 1024}" for i in range(8)] +
            [f"ds_read_b128 %[{b}{j}], %[rb{slot}] offset:{j * 1024}" for j in range(8)])


def mfmas(a, b, js):
    return [f"v_mfma_i32_16x16x64_i8 %[c{i}{j}], %[{b}{j}], %[{a}{i}], %[c{i}{j}]" for j in js for i in range(8)]


def interleave(mfma, extra, start, every):
    """The MFMA stream with (M0, load) pairs placed from MFMA `start` on, one pair per `every`
    MFMAs: M0 after one MFMA, its load after the next."""
    out, slots = [], {}
    for p, (m0, ld) in enumerate(extra):
        slots[start + p * every] = m0
        slots[start + p * every + 1] = ld
    for q, m in enumerate(mfma):
        out.append(m)
        if q in slots:
            out.append(slots[q])
    return out


def iteration(u, S, lbl_end):
    cur = ("fa", "fb") if u % 2 == 0 else ("ga", "gb")
    nxt = ("ga", "gb") if u % 2 == 0 else ("fa", "fb")
    out = ["s_waitcnt lgkmcnt(0)"]
    # the k-step it+S-1 loads into slot (u-1) % S: loads 0..3 in the first half, one per 8 MFMAs;
    # 4..7 in the second half after its LDS reads (a load costs the wave more issue time than an
    # MFMA gap holds, so they are spread thin: 1.04-1.05x over all 8 in the first half)
    pairs, advance = load_stage((u - 1) % S)
    out += interleave(mfmas(cur[0], cur[1], range(4)), pairs[:4], 0, 8)
    # half-way: k-step it+1 has landed for this wave (4 loads of k-step it+S-1 not yet issued),
    # then for the workgroup
    out.append(f"s_waitcnt vmcnt({8 * (S - 2) - 4})")
    out.append("s_barrier")
    # second half: MFMAs of B fragments 4..7, the reads of k-step it+1 after the first 16, the
    # other 4 loads after them
    mb = mfmas(cur[0], cur[1], range(4, 8))
    rd = reads((u + 1) % S, nxt[0], nxt[1])
    second = []
    for q in range(32):
        second.append(mb[q])
        if q < 16:
            second.append(rd[q])
    tail = interleave(second[32:], pairs[4:], 0, 4)  # second[32:] holds MFMAs 16..31
    out += second[:32] + tail + advance
    out += ["s_sub_u32 %[cnt], %[cnt], 1", "s_cmp_eq_u32 %[cnt], 0", f"s_cbranch_scc1 {lbl_end}"]
    return out


def generate(S):
    U = S if S % 2 == 0 else 2 * S  # iterations per unrolled block: whole ring turns, even
    lbl_loop, lbl_end = "L_crt_loop_%=", "L_crt_end_%="
    out = []
    # prologue: k-steps 0..S-2 into slots 0..S-2, wait for k-step 0, read it into F0
    for s in range(S - 1):
        pairs, advance = load_stage(s)
        for m0, ld in pairs:
            out += [m0, "s_nop 0", ld]
        out += advance
    out += [f"s_waitcnt vmcnt({8 * (S - 2)})", "s_barrier"] + reads(0, "fa", "fb")
    out.append(f"{lbl_loop}:")
    for u in range(U):
        out += iteration(u, S, lbl_end)
    out.append(f"s_branch {lbl_loop}")
    out.append(f"{lbl_end}:")
    # nothing may still land in LDS when the workgroup ends; MFMA results read by VALU next
    out += ["s_waitcnt vmcnt(0) lgkmcnt(0)", "s_nop 7", "s_nop 7", "s_nop 7"]
    return out


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    lines = ["// Generated by gen_crt_loop.py: the CRT residue-product main loop (see that script for the",
             "// schedule). Do not edit by hand.", "#pragma once", ""]
    for S in STAGES:
        body = generate(S)
        lines.append(f"#define TMM_CRT_LOOP_S{S} \\")
        lines += [f'  "{ins}\\n" \\' for ins in body]
        lines.append('  ""')
        lines.append("")
    # operand lists, in terms of the kernel's variables: acc[8][8], fa/fb/ga/gb[8], vo[4], ra/rb[S]
    outs = [f'[c{i}{j}] "+a"(acc[{i}][{j}])' for i in range(8) for j in range(8)]
    outs += [f'[{n}{i}] "=&v"({n}[{i}])' for n in ("fa", "fb", "ga", "gb") for i in range(8)]
    outs += ['[cnt] "+s"(cnt)', '[adv] "+s"(adv)', '[tmp] "=&s"(tmp)', '[tmp2] "=&s"(tmp2)']
    outs += [f'[vo{t}] "+v"(vo[{t}])' for t in range(4)]
    lines.append("#define TMM_CRT_OUTPUTS \\")
    lines += [f"  {o}, \\" for o in outs[:-1]] + [f"  {outs[-1]}", ""]
    for S in STAGES:
        ins = ['[ldsw] "s"(ldsw)', '[gA] "s"(gA)', '[gB] "s"(gB)']
        ins += [f'[r{x}{s}] "v"(r{x}[{s}])' for x in ("a", "b") for s in range(S)]
        lines.append(f"#define TMM_CRT_INPUTS_S{S} \\")
        lines += [f"  {o}, \\" for o in ins[:-1]] + [f"  {ins[-1]}", ""]
    text = "\n".join(lines)
    path = os.path.join(here, "tmm_crt_loop.inc")
    if len(sys.argv) > 1 and sys.argv[1] == "--check":
        return 0 if open(path).read() == text else 1
    with open(path, "w") as f:
        f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
