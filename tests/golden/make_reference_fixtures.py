"""Writes the reference's end-to-end known answers as fixtures.

Reads, from the reference repository (test/images/), the noisy inputs at
noise variance 0.05 and the clean references of its three test pictures, the
denoised outputs its GPU path saved for the published runs (8x8 patches,
256 atoms; 5 K-SVD iterations for barbara, 7 for cameraman and west), and the
published PSNRs of test/images/TEST.xlsx (sheets BARBARA, CAMERAMAN, WEST,
rows "256 ATOMI / n ITER" of the GPU table, column I = variance 0.05; sheet
RUMOROSE, column I = the noisy inputs).  Writes
  tests/golden/reference_images.npz  (uint8 pictures, compressed)
  tests/golden/reference_psnr.json   (the transcribed cells)
The spreadsheet is read as XML from the .xlsx archive; the PNGs with PIL.

Usage: python tests/golden/make_reference_fixtures.py [/path/to/reference]
"""
import json
import os
import re
import sys
import zipfile

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
IMG = os.path.join(REF, "test", "images")

# (picture, sheet, K-SVD iterations of the published GPU run, its row, the
# saved output's name: west's outputs leave the atom count out of the name)
CASES = [("barbara", "BARBARA", 5, 35, "8_8_barbara_0,05_256_5.png"),
         ("cameraman", "CAMERAMAN", 7, 37, "8_8_cameraman_0,05_256_7.png"),
         ("west", "WEST", 7, 37, "8_8_west_0,05_7.png")]
NOISY_ROW = {"barbara": 10, "cameraman": 8, "west": 12}  # RUMOROSE rows


def xlsx_cells(path):
    """{sheet name: {cell: value}} of every numeric or string cell."""
    z = zipfile.ZipFile(path)
    strings = ["".join(re.findall(r"<t[^>]*>(.*?)</t>", si, re.S))
               for si in re.findall(r"<si>(.*?)</si>", z.read("xl/sharedStrings.xml").decode(), re.S)]
    wb = z.read("xl/workbook.xml").decode()
    rels = z.read("xl/_rels/workbook.xml.rels").decode()
    target = dict(re.findall(r'Id="(rId\d+)"[^>]*Target="([^"]+)"', rels))
    target.update({a: b for b, a in re.findall(r'Target="([^"]+)"[^>]*Id="(rId\d+)"', rels)})
    out = {}
    for name, rid in re.findall(r'<sheet name="([^"]+)"[^>]*r:id="(rId\d+)"', wb):
        xml = z.read("xl/" + target[rid]).decode()
        cells = {}
        for ref, attrs, body in re.findall(r'<c r="([A-Z]+\d+)"([^>]*?)(?:/>|>(.*?)</c>)', xml, re.S):
            v = re.search(r"<v>(.*?)</v>", body or "")
            if v:
                cells[ref] = strings[int(v.group(1))] if 't="s"' in attrs else float(v.group(1))
        out[name] = cells
    return out


def load(path):
    a = np.asarray(Image.open(path))
    assert a.dtype == np.uint8 and a.ndim == 2, path
    return a


def main():
    cells = xlsx_cells(os.path.join(IMG, "TEST.xlsx"))
    arrays, psnr = {}, {"source": "test/images/TEST.xlsx", "variance": 0.05, "patch": [8, 8], "atoms": 256,
                        "cases": []}
    for pic, sheet, iters, row, outname in CASES:
        arrays[f"{pic}_noisy"] = load(os.path.join(IMG, "input", f"{pic}_0,05.png"))
        arrays[f"{pic}_ref"] = load(os.path.join(IMG, "input", f"{pic}_ref.png"))
        arrays[f"{pic}_gpu_out"] = load(os.path.join(IMG, "output", pic, "GPU", outname))
        label = cells[sheet][f"A{row}"]
        assert label == f"256 ATOMI / {iters} ITER", label
        psnr["cases"].append({"picture": pic, "ksvd_iter": iters, "gpu_psnr": cells[sheet][f"I{row}"],
                              "cell": f"{sheet}!I{row}", "noisy_psnr": cells["RUMOROSE"][f"I{NOISY_ROW[pic]}"],
                              "noisy_cell": f"RUMOROSE!I{NOISY_ROW[pic]}",
                              "output_file": f"test/images/output/{pic}/GPU/{outname}"})
    np.savez_compressed(os.path.join(HERE, "reference_images.npz"), **arrays)
    with open(os.path.join(HERE, "reference_psnr.json"), "w") as f:
        json.dump(psnr, f, indent=1)
    print("wrote", sorted(arrays), [c["gpu_psnr"] for c in psnr["cases"]])


if __name__ == "__main__":
    main()
