"""Copy the reference's own small fixture files for the resample / convolution
path into tests/golden/reference_fixtures/.

These are data files held by the reference's test suite (expected outputs of
deterministic closed-form inputs, most of them produced by libvips).  They are
copied verbatim as bytes; no reference source is copied.  Run once in a tree
where the reference checkout is available; the copies are committed so the
tests never read the reference at run time.

Usage: python tests/golden/import_reference_fixtures.py /path/to/viprs
"""
import pathlib
import shutil
import sys

DIRS = [
    "reduceh_libvips",
    "reducev_libvips",
    "reduce_libvips",
    "shrink_libvips",
    "shrinkh",
    "shrinkv",
    "gauss_blur",
    "gauss_blur_libvips",
    "convsep_libvips",
    "libvips_parity_resample",
    "libvips_parity_convolution",
    "thumbnail_libvips",
]


def main() -> None:
    ref = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    dst_root = pathlib.Path(__file__).resolve().parent / "reference_fixtures"
    for d in DIRS:
        src = ref / "tests" / "fixtures" / d
        dst = dst_root / d
        dst.mkdir(parents=True, exist_ok=True)
        for f in sorted(src.iterdir()):
            if f.suffix in (".bin", ".fmt"):
                shutil.copyfile(f, dst / f.name)
                print(f"{d}/{f.name} {f.stat().st_size}")


if __name__ == "__main__":
    main()
