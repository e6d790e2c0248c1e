"""Copy CP2K's grid known-answer files into tests/golden/sample_tasks/.

The 13 ``src/grid/sample_tasks/*.task`` files of the reference are data:
each holds one task's inputs (geometry, exponents, pab) and the reference
backend's outputs (collocated grid, integrated hab, forces, virial), in the
"#Grid task v10" text format read by src/grid/grid_replay.c:221-344.  They are
stored gzip-compressed, byte-identical after decompression.

Run once in a container that has the reference tree:
    python tests/golden/make_fixtures.py [/path/to/cp2k/src/grid/sample_tasks]
"""
import gzip
import pathlib
import sys

SRC = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else
                   "/root/reference/src/grid/sample_tasks")
DST = pathlib.Path(__file__).resolve().parent / "sample_tasks"


def main():
    DST.mkdir(exist_ok=True)
    files = sorted(SRC.glob("*.task"))
    if not files:
        sys.exit(f"no .task files under {SRC}")
    for f in files:
        data = f.read_bytes()
        with gzip.GzipFile(DST / (f.name + ".gz"), "wb", mtime=0) as out:
            out.write(data)
        print(f"{f.name}: {len(data)} bytes")


if __name__ == "__main__":
    main()
