"""Copy the reference's golden fixtures into tests/golden/fixtures as gzipped JSON.

The fixtures are data dumped from the Python original of embloks-core
(crates/embloks-core/tests/fixtures/*.json, read by
crates/embloks-core/tests/fixtures.rs): inputs and expected outputs only.
Run once in a checkout that has the reference at /root/reference.
"""
import gzip
import pathlib
import shutil
import sys

SOURCE = pathlib.Path("/root/reference/crates/embloks-core/tests/fixtures")
TARGET = pathlib.Path(__file__).resolve().parent / "fixtures"


def main() -> int:
    if not SOURCE.is_dir():
        print(f"no reference fixtures at {SOURCE}", file=sys.stderr)
        return 1
    TARGET.mkdir(exist_ok=True)
    for path in sorted(SOURCE.glob("*.json")):
        with path.open("rb") as src, gzip.open(TARGET / (path.name + ".gz"), "wb", compresslevel=9) as dst:
            shutil.copyfileobj(src, dst)
        print("copied", path.name)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
