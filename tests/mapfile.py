"""The reference's own scene fixture res/simple.qsf (a lightmapped QSF2 map),
kept gzip-compressed under tests/golden/, unpacked once per process for the
map workload (scene kind 4 reads its path from $SFTX_QSF)."""
import gzip
import os
import shutil
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = None


def map_path():
    global _PATH
    if _PATH is None:
        fd, path = tempfile.mkstemp(suffix=".qsf")
        with os.fdopen(fd, "wb") as out, gzip.open(os.path.join(HERE, "golden", "simple.qsf.gz"), "rb") as src:
            shutil.copyfileobj(src, out)
        _PATH = path
    os.environ["SFTX_QSF"] = _PATH
    return _PATH
