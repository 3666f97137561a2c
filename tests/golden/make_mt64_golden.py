"""Writes tests/golden/mt64_streams.json: the reference's own MtRng64::getReal2() stream
(src/vendor/mersennetwister.h, compiled by `make -C oracle ref` into oracle/_ref/mt_ref) for a few
seeds.  Run in a container that has the reference checkout."""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "mt_ref")
    out = {}
    for seed, n in [(0, 700), (1, 50), (12345678, 50), (18446744073709551615, 20)]:
        vals = subprocess.run([exe, str(seed), str(n)], check=True, capture_output=True, text=True).stdout.split()
        out[str(seed)] = [float(v) for v in vals]
    with open(os.path.join(HERE, "mt64_streams.json"), "w") as f:
        json.dump(out, f)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
