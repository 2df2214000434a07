"""Copy the reference crate's golden vectors for the nhswps hot path into tests/golden/ref.

The fixtures are data files held by the reference's own tests (big-endian f64, Fortran order,
no header; shapes encoded in the file name), plus the end-to-end case of src/test.rs:27-36 for
ng = 32, nz = 4: its balanced initial state (testdata/balinit/32_4_sw_init.r8, three little-endian
f64 records of an 8-byte header and ng^2 values: q, delta, gamma) and the run's outputs
(testdata/complete/32_4: monitor.asc, ecomp.asc and the .r4 field dumps), and the two balinit
cases of src/balinit.rs:441-481 (qq_init.r8 in, sw_init.r8 out, at 18x2 and 32x4).  This script copies them verbatim, converts the
insta snapshot values (factor lists and wavenumber arrays) into snapshots.json, and writes
MANIFEST.json with the SHA-256 and source path of every file.  It is the only script that reads
the reference tree; nothing at test time does.

Usage: python tests/golden/import_fixtures.py /root/reference
"""
import hashlib
import json
import os
import re
import shutil
import sys

DIRS = [
    "stafft/testdata/forfft", "stafft/testdata/revfft", "stafft/testdata/initfft",
    "testdata/init2dfft", "testdata/ptospc", "testdata/spctop", "testdata/deriv",
    "spectral/testdata/init_spectral", "spectral/testdata/main_invert", "spectral/testdata/jacob",
    "spectral/testdata/divs", "spectral/testdata/ptospc3d", "spectral/testdata/spctop3d",
    "spectral/testdata/deal3d", "spectral/testdata/deal2d", "nhswps/testdata/source",
    "nhswps/testdata/cpsource", "nhswps/testdata/vertical", "nhswps/testdata/coeffs",
    "nhswps/testdata/psolve", "nhswps/testdata/advance/18_6", "nhswps/testdata/advance/24_4",
]

# single files of the end-to-end case (src/test.rs:27-36)
EXTRA = ["testdata/balinit/32_4_sw_init.r8", "testdata/balinit/32_4_qq_init.r8",
         "testdata/balinit/18_2_qq_init.r8", "testdata/balinit/18_2_sw_init.r8", "testdata/complete/32_4/monitor.asc",
         "testdata/complete/32_4/ecomp.asc"] + \
        [f"testdata/complete/32_4/2d/{f}.r4" for f in ("d", "g", "h", "q", "zeta")] + \
        [f"testdata/complete/32_4/3d/{f}.r4" for f in ("d", "g", "pn", "ql", "r", "w")]


def snap_values(path):
    txt = open(path).read().split("---", 2)[2]
    return [float(x) if "." in x else int(x) for x in re.findall(r"-?[0-9][0-9.e+-]*", txt)]


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "ref")
    manifest = {}
    for d in DIRS:
        src = os.path.join(ref, "src", d)
        dst = os.path.join(out, d)
        os.makedirs(dst, exist_ok=True)
        for f in sorted(os.listdir(src)):
            if not f.endswith(".bin"):
                continue
            shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
            data = open(os.path.join(dst, f), "rb").read()
            manifest[f"{d}/{f}"] = {"sha256": hashlib.sha256(data).hexdigest(),
                                    "source": f"src/{d}/{f}", "bytes": len(data)}
    for rel in EXTRA:
        src = os.path.join(ref, "src", rel)
        dst = os.path.join(out, rel)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copyfile(src, dst)
        data = open(dst, "rb").read()
        manifest[rel] = {"sha256": hashlib.sha256(data).hexdigest(), "source": f"src/{rel}", "bytes": len(data)}
    snaps = {}
    for sd in ("src/stafft/snapshots", "src/snapshots"):
        for f in sorted(os.listdir(os.path.join(ref, sd))):
            snaps[f] = snap_values(os.path.join(ref, sd, f))
    with open(os.path.join(out, "snapshots.json"), "w") as fh:
        json.dump(snaps, fh, indent=0)
    with open(os.path.join(out, "MANIFEST.json"), "w") as fh:
        json.dump({"provenance": "rse-standrewscs/shallow-water, reference crate test fixtures "
                   "(no licence file ships with the reference)", "files": manifest}, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
