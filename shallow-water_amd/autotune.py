"""Autotuner of the block-per-rows transform passes (the transform lengths other than 128, 256 and 512,
which have wave-per-sequence kernels): for each transform length and value size it times every valid
tile -- rows per block B (B / 2 complex sequences of two packed real rows) and threads per block nt --
on a forward and an inverse 2-D transform of as many slabs as the workloads transform per launch
(sw_fft_time_tile), and writes the fastest to csrc/sw_fft_tiles.inc, which the library compiles in.
Results do not depend on the tile; only the speed does.  Run on the GPU, then rebuild:

    python shallow-water_amd/autotune.py [--lengths 64:100,1536:25] [--out shallow-water_amd/csrc/sw_fft_tiles.inc]
"""
import argparse
import importlib.util
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
# length: slabs per launch.  ng = 64: the ensemble workload's members (nz = 4 and 8, about eleven per batch,
# about 100 slabs per transform); ng = 96: a reference test length (nz = 6, batches of a few members);
# ng = 1536: the large mixed-radix case (nz = 24, 25 slabs)
DEFAULT = "64:100,96:56,1536:25"


def load():
    spec = importlib.util.spec_from_file_location("shallow_water_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["shallow_water_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def candidates(n):
    bs = [b for b in range(2, min(n, 64) + 1, 2) if n % b == 0] if n % 2 == 0 else [1]
    return [(b, nt) for b in bs for nt in range(64, 513, 64)]


def measure(sw, n, nbytes, b, nt, nslab, target_ms=25.0):
    us = sw.fft_time_tile(n, nbytes, b, nt, nslab, 2)
    reps = max(3, min(200, int(target_ms * 1e3 / max(us, 1.0))))
    return min(sw.fft_time_tile(n, nbytes, b, nt, nslab, reps) for _ in range(2))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--lengths", default=DEFAULT, help="n:slabs,... (default %(default)s)")
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "sw_fft_tiles.inc"))
    args = ap.parse_args()
    sw = load()
    sw.init(0)
    rows, notes = [], []
    for item in args.lengths.split(","):
        n, nslab = (int(x) for x in item.split(":"))
        for nbytes in (8, 4):
            sw.fft_set_tile(n, nbytes, 0, 0)
            rule = sw.fft_get_tile(n, nbytes)
            t0 = time.time()
            res = []
            for b, nt in candidates(n):
                try:
                    res.append((measure(sw, n, nbytes, b, nt, nslab), b, nt))
                except sw.SwError:
                    continue  # not a valid tile for this length
            res.sort()
            best = res[0]
            t_rule = next(t for t, b, nt in res if (b, nt) == rule)
            line = (f"n={n} {'FP64' if nbytes == 8 else 'FP32'} {nslab} slabs: best B={best[1]} nt={best[2]} "
                    f"{best[0]:.1f} us, rule B={rule[0]} nt={rule[1]} {t_rule:.1f} us ({t_rule / best[0]:.2f}x), "
                    f"{len(res)} tiles in {time.time() - t0:.1f} s")
            print(line, flush=True)
            notes.append(line)
            rows.append((n, nbytes, best[1], best[2]))
    with open(args.out, "w") as f:
        f.write("// Measured tiles of the block-per-rows transform passes, {n, value bytes, rows per block, threads}:\n")
        f.write("// the fastest forward + inverse 2-D transform over every valid tile (autotune.py, on one MI355X).\n")
        for line in notes:
            f.write(f"//   {line}\n")
        f.write("static const FftTile kFftTiles[] = {\n")
        for n, nbytes, b, nt in rows:
            f.write(f"    {{{n}, {nbytes}, {b}, {nt}}},\n")
        f.write("};\n")
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
