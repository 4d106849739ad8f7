"""`BP 3F M` with a lost or damaged shard (bp3f.f:646-685): the merge reads
and checks every partial volume before any device work, so a missing X or
_W file, a truncated file or a shard of another size stops it with an error
and no output volume.  (The reference stops at the file it cannot open;
a weights stack that ends early is its error 101.)  Runs on the CPU: the
failures come before the device is touched."""
import os

import numpy as np
import pytest

from spider_amd import spiderio


def _shards(d, nx=6, n=3, seed=1):
    rng = np.random.default_rng(seed)
    paths = []
    for k in range(n):
        X = (rng.standard_normal((2 * nx, 2 * nx, nx + 1)) + 1j * rng.standard_normal((2 * nx, 2 * nx, nx + 1)))
        W = rng.uniform(0, 2, (2 * nx, 2 * nx, nx + 1))
        p = os.path.join(d, f"part{k + 1}.dat")
        spiderio.write_partial(p, X.astype(np.complex64), W.astype(np.float32))
        paths.append(p)
    return paths


def test_lost_weights_shard(tmp_path):
    from spider_amd import bp
    paths = _shards(str(tmp_path))
    os.remove(spiderio._weights_name(paths[1]))
    out = str(tmp_path / "merged.dat")
    with pytest.raises(FileNotFoundError, match="part2_W"):
        bp.bp3f_merge_files(paths, out)
    assert not os.path.exists(out)


def test_lost_volume_shard(tmp_path):
    from spider_amd import bp
    paths = _shards(str(tmp_path))
    os.remove(paths[2])
    out = str(tmp_path / "merged.dat")
    with pytest.raises(FileNotFoundError, match="part3"):
        bp.bp3f_merge_files(paths, out)
    assert not os.path.exists(out)


def test_truncated_shard(tmp_path):
    from spider_amd import bp
    paths = _shards(str(tmp_path))
    size = os.path.getsize(paths[0])
    with open(paths[0], "r+b") as fh:
        fh.truncate(size - 100)
    out = str(tmp_path / "merged.dat")
    with pytest.raises(spiderio.SpiderFormatError, match="truncated"):
        bp.bp3f_merge_files(paths, out)
    assert not os.path.exists(out)


def test_shard_of_another_size(tmp_path):
    from spider_amd import bp
    paths = _shards(str(tmp_path))
    os.makedirs(tmp_path / "other")
    odd = _shards(str(tmp_path / "other"), nx=8, n=1)
    out = str(tmp_path / "merged.dat")
    with pytest.raises(spiderio.SpiderFormatError, match="differs"):
        bp.bp3f_merge_files(paths + odd, out)
    assert not os.path.exists(out)
