#!/usr/bin/env python3
"""Monte Carlo LDPC sweep driver on an MI355X -- drop-in for the reference's
run_simulation.py (same flags, console table and CSV; run_simulation.py:59-256).

By default the engine runs in this process: libldpc through ctypes
(ldpc_mi355x.decoder.Engine), Profile R -- the scalar_5w arithmetic the
reference driver runs -- on the given matrix.  ``--engine circulant`` runs the
golden circulant arithmetic (Profile T, AR4JA r1/2 K=1024) and writes the
circulant engine's 9-column CSV.  ``--binary PATH`` keeps the reference's
process protocol instead: argv ``-c -e -p -t -k -i [-s] [-n] -b [-L -R -V]``
and stdout JSON lines (progress lines, then one final line); our own
``ldpc_mi355x/ldpc_sim_scalar`` speaks it.

Block accounting follows the reference: ``--max_blocks_per_core`` counts per
core of a 440-core mesh (1 with ``--single_core``), ``--max_blocks`` is a
total, ``--batch_size`` is per core between stopping-rule checks.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

from ldpc_mi355x import codes  # noqa: E402

parse_matrix_header = codes.parse_matrix_header
parse_chinn_header = parse_matrix_header

CSV_HEADER = "EbN0_dB,Total_Blocks,Total_Bit_Errors,Total_Block_Errors,BER,FER\n"
CSV_HEADER_CIRCULANT = ("EbN0_dB,Total_Blocks,Total_Bit_Errors,Total_Block_Errors,BER,FER,Mean_Iterations,"
                        "Throughput_Msps,Host_Wall_s\n")
CORES_MESH = 440


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="MI355X LDPC Monte Carlo Simulator Driver")
    p.add_argument("--chinn", "--matrix", dest="chinn", type=str, required=True,
                   help="parity-check matrix: a .chinn.out triplet file or a legacy .chinn")
    p.add_argument("--ebn0_start", type=float, required=True, help="first Eb/N0 point (dB)")
    p.add_argument("--ebn0_end", type=float, required=True, help="last Eb/N0 point (dB), included")
    p.add_argument("--ebn0_step", type=float, default=0.25, help="spacing of the Eb/N0 points (dB)")
    p.add_argument("--punctured", type=int, default=0, help="how many trailing variables are punctured (never transmitted)")
    p.add_argument("--min_errors", "--target_errors", dest="min_errors", type=int, default=50,
                   help="stop a point once this many frame errors are counted (default 50)")
    p.add_argument("--max_blocks", type=int, default=None,
                   help="cap on codewords per point; takes precedence over --max_blocks_per_core")
    p.add_argument("--max_blocks_per_core", "--blocks_per_core", dest="max_blocks_per_core", type=int,
                   default=100000, help="cap on codewords per point, per reference core (x 440, or x 1 with --single_core)")
    p.add_argument("--batch_size", type=int, default=1000,
                   help="codewords per reference core between progress reports and stop checks (default 1000)")
    p.add_argument("--max_iter", "--max_iterations", dest="max_iter", type=int, default=16,
                   help="iteration cap of the decoder (default 16)")
    p.add_argument("--binary", type=str, default=None,
                   help="Engine binary speaking the JSON-lines protocol (default: in-process libldpc)")
    p.add_argument("--output", type=str, default=None,
                   help="CSV to write (default: a name built from the run's parameters)")
    p.add_argument("--single_core", action="store_true", help="Block accounting for one core instead of 440")
    p.add_argument("--nms", action="store_true",
                   help="decode with normalized min-sum (scale 0.75) rather than the recursive Approximate-Min*")
    p.add_argument("--l_max", type=float, default=0.0, help="clip channel LLRs to this magnitude (0: no clip)")
    p.add_argument("--r_max", type=float, default=0.0, help="clip check-to-variable messages to this magnitude (0: no clip)")
    p.add_argument("--l_post_max", "--v_max", dest="l_post_max", type=float, default=0.0,
                   help="clip the updated variable LLRs to this magnitude (0: no clip)")
    p.add_argument("--engine", type=str, default="scalar", choices=["scalar", "circulant"],
                   help="'scalar': the scalar_5w arithmetic on the matrix (Profile R); 'circulant': the golden "
                        "circulant arithmetic (Profile T, AR4JA r1/2 K=1024)")
    p.add_argument("--seed", type=int, default=133742, help="Philox key of the channel stream (default: 133742)")
    p.add_argument("--beta", type=float, default=0.9375, help="Profile T damping (circulant engine)")
    p.add_argument("--device", type=int, default=0, help="HIP device index")
    p.add_argument("--random_info", action="store_true",
                   help="circulant engine: send random information bits, encoded, instead of the all-zero "
                        "codeword, and count errors against the sent bits")
    return p


def output_filename(args) -> str:
    """The reference's CSV path rule (run_simulation.py:108-137)."""
    results_dir = "Results"
    if args.output:
        if os.path.dirname(args.output):
            os.makedirs(os.path.dirname(args.output), exist_ok=True)
            return args.output
        return os.path.join(results_dir, args.output)
    base = os.path.basename(args.chinn)
    for ext in (".chinn.out", ".chinn", ".out"):
        if base.endswith(ext):
            base = base[:-len(ext)]
            break
    algo = "nms" if args.nms else "amin"
    core_str = "1core" if args.single_core else "440cores"
    clip = []
    if args.l_max > 0.0:
        clip.append(f"L{args.l_max:.1f}")
    if args.r_max > 0.0:
        clip.append(f"R{args.r_max:.1f}")
    if args.l_post_max > 0.0:
        clip.append(f"V{args.l_post_max:.1f}")
    clip_str = f"_clip{''.join(clip)}" if clip else ""
    return os.path.join(
        results_dir,
        f"results_{base}_{algo}{'_' + args.engine if args.engine != 'scalar' else ''}{clip_str}"
        f"_ebn0_{args.ebn0_start:.2f}_{args.ebn0_end:.2f}_step{args.ebn0_step:.2f}"
        f"_max{args.max_blocks_per_core}_min{args.min_errors}_iter{args.max_iter}_{core_str}.csv")


def csv_row(ebn0, blocks, bit_errors, frame_errors, unpunctured_n) -> str:
    total = blocks * unpunctured_n
    ber = bit_errors / total if total > 0 else 0.0
    fer = frame_errors / blocks if blocks > 0 else 0.0
    return f"{ebn0:.2f},{blocks},{bit_errors},{frame_errors},{ber:.6e},{fer:.6e}\n"


def csv_row_circulant(ebn0, blocks, bit_errors, frame_errors, iter_sum, seconds) -> str:
    """ldpc_sim_circulant_dma.cpp:395-406: BER over 1024 info bits, Msps over 2048 symbols."""
    ber = bit_errors / (blocks * 1024.0) if blocks else 0.0
    fer = frame_errors / blocks if blocks else 0.0
    mi = iter_sum / blocks if blocks else 0.0
    msps = blocks * 2048.0 / (seconds * 1e6) if seconds > 0 else 0.0
    return f"{ebn0:.2f},{blocks},{bit_errors},{frame_errors},{ber:.6e},{fer:.6e},{mi:.3f},{msps:.2f},{seconds:.2f}\n"


def progress_line(ebn0, d) -> str:
    pct = (d["blocks"] / d["max_blocks"] * 100.0) if d.get("max_blocks", 0) > 0 else 0.0
    return (f"  [{ebn0:.2f} dB] Blocks: {d['blocks']:,}/{d['max_blocks']:,} ({pct:5.1f}%) "
            f"| FE: {d['frame_errors']}/{d['target_errors']} | FER: {d['fer']:.2e} | BER: {d['ber']:.2e} "
            f"| {d['throughput_msps']:.1f} Msps | {d['elapsed_sec']:.1f}s")


def run_point_binary(args, ebn0, is_tty):
    """The reference's subprocess protocol (run_simulation.py:165-225)."""
    cmd = [args.binary, "-c", args.chinn, "-e", str(ebn0), "-p", str(args.punctured), "-t", str(args.min_errors),
           "-k", str(args.max_blocks_per_core), "-i", str(args.max_iter)]
    if args.single_core:
        cmd.append("-s")
    if args.nms:
        cmd.append("-n")
    if args.batch_size is not None:
        cmd.extend(["-b", str(args.batch_size)])
    if args.l_max > 0.0:
        cmd.extend(["-L", str(args.l_max)])
    if args.r_max > 0.0:
        cmd.extend(["-R", str(args.r_max)])
    if args.l_post_max > 0.0:
        cmd.extend(["-V", str(args.l_post_max)])
    final = None
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, bufsize=1)
    for raw in proc.stdout:
        line = raw.strip()
        if not line:
            continue
        if line.startswith("{") and "ebn0" in line:
            try:
                data = json.loads(line)
            except ValueError:
                continue
            if data.get("progress"):
                s = progress_line(ebn0, data)
                print(f"\r{s:<104}", end="", flush=True) if is_tty else print(s, flush=True)
            else:
                final = data
        else:
            print(f"\n{line}" if is_tty else line)
    proc.wait()
    if proc.returncode != 0:
        raise RuntimeError(f"Simulation process exited with code {proc.returncode}")
    if final is None:
        raise RuntimeError("Could not parse final simulation output JSON.")
    return final["blocks"], final["bit_errors"], final["frame_errors"], final.get("elapsed_sec", 0.0), None


def run_point_inprocess(args, engine, ebn0, unpunctured_n, is_tty):
    """Same stopping rule on the in-process engine: batches of batch_size x
    cores codewords until min_errors frame errors or the block budget."""
    cores = 1 if args.single_core else CORES_MESH
    max_blocks = args.max_blocks_per_core * cores
    batch = min(args.batch_size, args.max_blocks_per_core) * cores
    acc = None
    first = 0
    t0 = time.perf_counter()
    while (acc is None or acc.frame_errors < args.min_errors) and (acc is None or acc.blocks < max_blocks):
        n = min(batch, max_blocks - (acc.blocks if acc else 0))
        acc = engine.run_point(ebn0, n, seed=args.seed, first_cw=first, chunk_cw=min(n, 1 << 22), result=acc)
        first = acc.next_cw
        el = time.perf_counter() - t0
        d = dict(blocks=acc.blocks, max_blocks=max_blocks, frame_errors=acc.frame_errors,
                 target_errors=args.min_errors, fer=acc.fer,
                 ber=acc.bit_errors / (acc.blocks * unpunctured_n) if acc.blocks else 0.0,
                 throughput_msps=acc.blocks * unpunctured_n / el / 1e6 if el > 0 else 0.0, elapsed_sec=el)
        s = progress_line(ebn0, d)
        print(f"\r{s:<104}", end="", flush=True) if is_tty else print(s, flush=True)
    return acc.blocks, acc.bit_errors, acc.frame_errors, time.perf_counter() - t0, acc


def make_engine(args, N, M):
    from ldpc_mi355x import decoder
    if args.engine == "circulant":
        if (N, M, args.punctured) != (2560, 1536, 512):
            raise SystemExit("--engine circulant decodes the AR4JA r1/2 K=1024 code (N=2560, M=1536, 512 punctured)")
        eng = decoder.Engine(1024, "1/2", max_iter=args.max_iter, beta=args.beta, profile="T", device=args.device)
        if args.random_info:
            eng.set_codeword_mode(True)
        return eng
    if args.random_info:
        raise SystemExit("--random_info is built for the circulant engine (Profile T, AR4JA r1/2 K=1024)")
    with open(args.chinn) as f:
        first = next((ln for ln in f if ln.strip() and not ln.lstrip().startswith("#")), "")
    rows = codes.read_chinn_out(args.chinn)[2] if len(first.split()) == 3 else read_chinn_rows(args.chinn)
    return decoder.Engine(matrix=(rows, N), n_punct=args.punctured, max_iter=args.max_iter,
                          profile="R_NMS" if args.nms else "R", l_max=args.l_max, r_max=args.r_max,
                          l_post_max=args.l_post_max, device=args.device)


def read_chinn_rows(path):
    """Rows of a legacy .chinn file: header "N M [dv dc]", then per row "deg v1 .." (1-based)."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
    hdr = lines[0]
    m = int(hdr[1])
    body = lines[1:] if len(hdr) >= 4 else lines[2:]
    return [[int(v) - 1 for v in r[1:1 + int(r[0])]] for r in body[:m]]


def main(argv=None):
    args = build_parser().parse_args(argv)
    if args.random_info and args.binary:
        raise SystemExit("--random_info runs in process (the engines' argv protocol has no such flag)")
    total_cores = 1 if args.single_core else CORES_MESH
    if args.max_blocks is not None:
        args.max_blocks_per_core = max(1, args.max_blocks // total_cores)
    N, M, max_vdeg, max_cdeg = parse_matrix_header(args.chinn)
    unpunctured_n = N - args.punctured
    rate = (N - M) / unpunctured_n
    os.makedirs("Results", exist_ok=True)
    out_name = output_filename(args)
    circ = args.engine == "circulant" and args.binary is None
    algo_desc = "Normalized Min-Sum (alpha=0.75)" if args.nms else "Approximate-Min* (Christopher Jones MILCOM 2003)"
    if args.engine == "circulant":
        algo_desc = f"Tree Approximate-Min* (golden circulant arithmetic, beta={args.beta})"
        if args.random_info:
            algo_desc += ", random info bits"
    print("==========================================================")
    print("         MI355X LDPC Monte Carlo Simulation               ")
    print("==========================================================")
    print(f"Matrix File        : {args.chinn}")
    print(f"Decoder Algorithm  : {algo_desc}")
    print(f"Decoder Engine     : {args.engine.upper()} ({'process ' + args.binary if args.binary else 'libldpc'})")
    print(f"Code Dimensions    : N={N}, M={M} (Rate = {rate:.4f})")
    print(f"Punctured Nodes    : {args.punctured} (Transmitted N = {unpunctured_n} symbols/code)")
    print(f"Eb/N0 Range        : {args.ebn0_start:.2f} dB to {args.ebn0_end:.2f} dB (step {args.ebn0_step:.2f} dB)")
    print(f"Stopping Criteria  : Min {args.min_errors} block errors OR max {args.max_blocks_per_core} blocks/core")
    print(f"Batch Size/Core    : {args.batch_size} blocks per core per evaluation step")
    print(f"Max Decoder Iters  : {args.max_iter}")
    if args.l_max > 0.0 or args.r_max > 0.0 or args.l_post_max > 0.0:
        print(f"LLR Clipping       : L_max={args.l_max:.2f}, R_max={args.r_max:.2f}, L_post_max={args.l_post_max:.2f}")
    print(f"Output File        : {out_name}")
    print("==========================================================\n")

    points = np.arange(args.ebn0_start, args.ebn0_end + args.ebn0_step / 2.0, args.ebn0_step)
    with open(out_name, "w") as f:
        f.write(CSV_HEADER_CIRCULANT if circ else CSV_HEADER)
    print(f"{'Eb/N0 (dB)':<10} | {'Blocks Run':<12} | {'Bit Errors':<12} | {'Block Errors':<14} | {'BER':<12} | "
          f"{'FER':<12} | {'Throughput (Msps)':<18}")
    print("-" * 104)
    is_tty = sys.stdout.isatty()
    engine = None
    try:
        for ebn0 in points:
            try:
                if args.binary:
                    blocks, be, fe, sec, acc = run_point_binary(args, float(ebn0), is_tty)
                else:
                    if engine is None:
                        engine = make_engine(args, N, M)
                    blocks, be, fe, sec, acc = run_point_inprocess(args, engine, float(ebn0), unpunctured_n, is_tty)
                total = blocks * unpunctured_n
                ber = be / total if total > 0 else 0.0
                fer = fe / blocks if blocks > 0 else 0.0
                msps = total / sec / 1e6 if sec > 0 else 0.0
                row = (f"{ebn0:<10.2f} | {blocks:<12d} | {be:<12d} | {fe:<14d} | {ber:<12.4e} | {fer:<12.4e} | "
                       f"{msps:<18.2f}")
                print(f"\r{row:<104}" if is_tty else f"-> Final: {row}")
                with open(out_name, "a") as f:
                    if circ:
                        f.write(csv_row_circulant(ebn0, blocks, be, fe, acc.iter_sum, sec))
                    else:
                        f.write(csv_row(ebn0, blocks, be, fe, unpunctured_n))
            except Exception as e:  # the reference stops the sweep at the first failing point
                print(f"\nError simulating Eb/N0 = {ebn0:.2f} dB: {e}", file=sys.stderr)
                break
    finally:
        if engine is not None:
            engine.close()
    print(f"\nSimulation complete. Output saved to: {out_name}")
    return out_name


if __name__ == "__main__":
    main()
