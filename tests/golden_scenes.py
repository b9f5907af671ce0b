"""Scene recipes shared by the golden-fixture writer and the tests that read the fixtures."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

from make_golden import RENDERS, make_scene, soup_rays, soup_scene  # noqa: E402,F401
