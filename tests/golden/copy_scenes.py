"""Copy the reference's scene files (Liquid2dRender/scenes/*.json, data in
the reference's scene format, README.md:34-135) into tests/golden/scenes as
fixtures for the JSON scene reader's tests.  Run once where the reference
tree exists; the copies are committed."""
import os
import shutil

SRC = "/root/reference/Liquid2dRender/scenes"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes")
NAMES = ["waterfall.json", "fountain.json", "glass_pour.json", "level_eq.json", "dam_break.json", "drop.json",
         "double_dam_break.json", "test_scene.json", "viscosity_test.json", "smoke_test.json"]

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for n in NAMES:
        shutil.copyfile(os.path.join(SRC, n), os.path.join(DST, n))
        print("copied", n)
