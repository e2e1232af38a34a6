"""GPU-aware PIV topology: a series of image files through device batches.

Counterpart of src/fluidimage/topologies/piv.py:24-260 (``TopologyPIV``)
for one GPU. The reference's topology is a graph of works run by an
executor over CPU processes: names -> paths -> "read array" (io) ->
"make couples of arrays" -> "compute piv" (one couple per call) -> "save
piv" (io). Here the couples of the series are computed in batches of
``params.gpu.batch_pairs`` same-shape couples, each batch one run of the
batched GPU pipeline (every window of every couple of a pass in one
launch), while a pool of ``params.gpu.nb_io_workers`` threads reads the
images of the next batches and another thread saves the results of the
previous one:

    read (thread pool, PIL) -> [batch k + 1, k + 2 read ahead]
    compute (GPU, batch k)  -> save (thread, batch k - 1)

The series, the result names and the saving modes follow the reference:

* ``params.series``: ``path`` (a directory of images, or a file pattern
  with ``*``), ``str_subset`` (``"i:i+2"`` by default for numbered files:
  couples of images i and i + 1; ``"i, 0:2"`` for files numbered with a
  letter, ``c060a`` / ``c060b``: the two letters of one number), and
  ``ind_start``, ``ind_stop``, ``ind_step`` over i (fluiddyn's
  SeriesOfArrays, used at topologies/piv.py:82-88);
* ``params.saving``: ``path`` (default: the image directory + "." +
  ``postfix``), ``how`` in ("ask", "new_dir", "complete", "recompute")
  with the meaning of prepare_path_dir_result (topologies/base.py:49-104;
  "ask" without a terminal acts as "complete" when the directory exists);
  in "complete" mode the couples whose result file exists are skipped
  (compute_indices_to_be_computed, topologies/piv.py:160-167);
* result names ``piv_<first>-<last>`` (get_name_piv,
  data_objects/piv.py:37-57).

The results are written as NumPy ``.npz`` archives (the reference writes
HDF5 through h5py, which this environment lacks): per pass ``piv<k>/``
``deltaxs``, ``deltays``, ``xs``, ``ys``, ``correls_max``, the error
explanations as two arrays (``ierrors``, ``errors``), and the final
interpolated fields of the last pass.
"""

import glob
import json
import logging
import os
import re
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

from ..params import Params
from ..works.piv.multipass import WorkPIV

logger = logging.getLogger("fluidimage_amd.topologies")

IMAGE_EXTENSIONS = (".png", ".tif", ".tiff", ".bmp", ".jpg", ".jpeg", ".pgm")
HOW_VALUES = ("ask", "new_dir", "complete", "recompute")


def imread(path):
    """An image file as a 2-D array of its own type (uint8, uint16 or
    float32), as fluidimage.util.imread reads it."""
    from PIL import Image

    with Image.open(path) as im:
        if im.mode in ("I;16", "I;16B", "I;16L"):
            arr = np.array(im, dtype=np.uint16)
        elif im.mode == "F":
            arr = np.array(im, dtype=np.float32)
        elif im.mode == "I":
            arr = np.array(im)
            arr = arr.astype(np.uint16) if arr.min() >= 0 and arr.max() < 65536 else \
                arr.astype(np.float32)
        else:
            arr = np.array(im.convert("L"))
    return arr


_NAME = re.compile(r"^(?P<base>.*?)(?P<num>\d+)(?P<letter>[a-zA-Z]?)$")


class SeriesOfCouples:
    """The couples of a series of image files: [(name0, name1), ...] with
    their result names, from the file names and ``str_subset``."""

    def __init__(self, path, str_subset=None, ind_start=0, ind_stop=None, ind_step=1):
        path = str(path)
        if os.path.isdir(path):
            files = [f for f in sorted(os.listdir(path)) if f.lower().endswith(IMAGE_EXTENSIONS)]
            self.path_dir = Path(path)
        else:
            files = sorted(os.path.basename(f) for f in glob.glob(path))
            self.path_dir = Path(os.path.dirname(path) or ".")
        entries = []
        for f in files:
            stem = os.path.splitext(f)[0]
            m = _NAME.match(stem)
            if m:
                entries.append((m.group("base"), m.group("num"), m.group("letter"), f))
        if not entries:
            raise ValueError(f"no numbered image files in {path}")
        lettered = any(e[2] for e in entries)
        if str_subset is None:
            str_subset = "i, 0:2" if lettered else "i:i+2"
        self.str_subset = str_subset
        self.couples = []
        self.names = []
        if lettered:
            self._couples_lettered(entries, str_subset, ind_start, ind_stop, ind_step)
        else:
            self._couples_numbered(entries, str_subset, ind_start, ind_stop, ind_step)

    @staticmethod
    def _indices(values, ind_start, ind_stop, ind_step):
        lo = values[0] if ind_start in (None, "first") else int(ind_start)
        hi = values[-1] + 1 if ind_stop is None else int(ind_stop)
        return range(lo, hi, int(ind_step or 1))

    def _couples_numbered(self, entries, str_subset, ind_start, ind_stop, ind_step):
        # "i+a:i+b[:s]": the images range(i + a, i + b, s), two of them
        m = re.fullmatch(r"\s*i\s*([+-]\s*\d+)?\s*:\s*i\s*([+-]\s*\d+)?\s*(?::\s*(\d+))?\s*",
                         str_subset)
        if m is None:
            raise NotImplementedError(f"str_subset {str_subset!r} (use 'i+a:i+b[:step]')")
        a = int(m.group(1).replace(" ", "")) if m.group(1) else 0
        b = int(m.group(2).replace(" ", "")) if m.group(2) else 0
        s = int(m.group(3)) if m.group(3) else 1
        if len(range(a, b, s)) != 2:
            raise ValueError(f"str_subset {str_subset!r} does not select two images")
        by_num = {int(num): (num, f) for _, num, _, f in entries}
        nums = sorted(by_num)
        for i in self._indices(nums, ind_start, ind_stop, ind_step):
            k0, k1 = i + a, i + a + s
            if k0 in by_num and k1 in by_num:
                (s0, f0), (s1, f1) = by_num[k0], by_num[k1]
                self.couples.append((f0, f1))
                self.names.append(f"piv_{s0}-{s1}")

    def _couples_lettered(self, entries, str_subset, ind_start, ind_stop, ind_step):
        # "i, 0:2": number i, letters 0 and 1 of that number
        m = re.fullmatch(r"\s*i\s*,\s*(\d+)\s*:\s*(\d+)\s*", str_subset)
        if m is None or int(m.group(2)) - int(m.group(1)) != 2:
            raise NotImplementedError(f"str_subset {str_subset!r} (use 'i, a:a+2')")
        l0 = int(m.group(1))
        by_num = {}
        for _, num, letter, f in entries:
            by_num.setdefault(int(num), (num, {}))[1][letter] = f
        nums = sorted(by_num)
        for i in self._indices(nums, ind_start, ind_stop, ind_step):
            if i not in by_num:
                continue
            s, letters = by_num[i]
            keys = sorted(letters)
            if len(keys) < l0 + 2:
                continue
            c0, c1 = keys[l0], keys[l0 + 1]
            self.couples.append((letters[c0], letters[c1]))
            self.names.append(f"piv_{s}{c0}{c1}")

    def __len__(self):
        return len(self.couples)


def prepare_path_dir_result(path_dir_input, path_saving, postfix, how):
    """topologies/base.py:49-104 without the interactive question."""
    if how not in HOW_VALUES:
        raise ValueError(f"how_saving (here equal to '{how}') should be in {HOW_VALUES}")
    path = str(path_saving) if path_saving is not None else str(path_dir_input) + "." + postfix
    if not os.path.exists(path):
        how = "recompute" if how == "ask" else how
    else:
        if how == "ask":
            how = "complete"
        if how == "new_dir":
            i = 0
            while os.path.exists(path + str(i)):
                i += 1
            path += str(i)
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    return path, how


def save_results(path, results, names, params):
    """One couple's multipass results as an .npz archive (written to a
    temporary name, then renamed, so that "complete" never sees a partial
    file)."""
    arrays = {"names": np.array(names), "params": np.array(json.dumps(params, default=str))}
    for k, r in enumerate(results.passes):
        p = f"piv{k}/"
        for key in ("deltaxs", "deltays", "xs", "ys", "correls_max"):
            v = getattr(r, key, None)
            if v is not None:
                arrays[p + key] = np.asarray(v)
        errs = getattr(r, "errors", None) or {}
        arrays[p + "ierrors"] = np.array(sorted(errs), dtype=np.int64)
        arrays[p + "errors"] = np.array([errs[i] for i in sorted(errs)], dtype=str)
        for key in ("deltaxs_final", "deltays_final", "ixvecs_final", "iyvecs_final"):
            v = getattr(r, key, None)
            if v is not None:
                arrays[p + key] = np.asarray(v)
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "wb") as f:
        np.savez(f, **arrays)
    os.replace(tmp, path)
    return path


class TopologyPIV:
    """PIV over a series of image files, batched on the GPU."""

    _short_name = "piv"
    WorkVelocimetry = WorkPIV

    @classmethod
    def create_default_params(cls):
        """topologies/piv.py:57-77: the saving parameters, the WorkPIV tree
        (series, piv0, fix, multipass, mask) and, for this topology, the
        batching of the GPU (no counterpart in the reference)."""
        params = Params(tag="params")
        params._set_child("saving", attribs={"path": None, "how": "ask",
                                             "postfix": cls._short_name})
        cls.WorkVelocimetry._complete_params_with_default(params)
        params._set_child("preproc", attribs={"im2im": None, "args_im2im": None})
        params._set_child("gpu", attribs={"batch_pairs": 16, "nb_io_workers": 4,
                                          "read_ahead_batches": 2})
        return params

    def __init__(self, params, logging_level="info", nb_max_workers=None):
        self.params = params
        if logging_level:
            logging.basicConfig()
            logger.setLevel(getattr(logging, str(logging_level).upper(), logging.INFO))
        s = params.series
        self.series = SeriesOfCouples(s.path, s.str_subset, s.ind_start, s.ind_stop, s.ind_step)
        self.path_dir_src = self.series.path_dir
        self.path_dir_result, self.how_saving = prepare_path_dir_result(
            self.path_dir_src, params.saving.path, params.saving.postfix, params.saving.how)
        self.nb_max_workers = nb_max_workers
        self.work_piv = None  # made at the first compute (needs the GPU)
        self.results = []
        self.durations = {"read": [], "compute_piv": [], "save_piv": []}
        self.errors = {}

    def compute_indices_to_be_computed(self):
        """topologies/piv.py:160-167: the couples without a result file."""
        return [k for k, name in enumerate(self.series.names)
                if not (self.path_dir_result / (name + ".npz")).exists()]

    def _im2im(self):
        p = self.params.preproc
        if p.im2im is None:
            return None
        from ..works.preproc import toolbox

        name = str(p.im2im).rsplit(".", 1)[-1]
        if name not in toolbox.__all__:
            raise ValueError(f"preproc.im2im {p.im2im!r}: one of {toolbox.__all__}")
        fn = getattr(toolbox, name)
        kwargs = dict(p.args_im2im or {})
        return lambda img: fn(img, **kwargs)

    def _read(self, k):
        t0 = time.perf_counter()
        f0, f1 = self.series.couples[k]
        a = imread(self.path_dir_src / f0)
        b = imread(self.path_dir_src / f1)
        self.durations["read"].append(time.perf_counter() - t0)
        return k, a, b

    def compute(self, executor=None, nb_max_workers=None, sleep_time=None,
                sequential=None, stop_if_error=False):
        """Compute the couples of the series (those without a result in
        "complete" mode). ``executor`` and the CPU-worker arguments of the
        reference are accepted; the work runs on the GPU."""
        if executor not in (None, "gpu", "exec_async", "exec_sequential",
                            "multi_exec_async", "exec_async_sequential"):
            raise ValueError(f"unknown executor {executor!r}")
        t_start = time.perf_counter()
        indices = (self.compute_indices_to_be_computed() if self.how_saving == "complete"
                   else list(range(len(self.series))))
        if not indices:
            logger.warning('topology in mode "complete" and work already done.'
                           if self.how_saving == "complete" else "Nothing to do")
            return
        logger.info("Add %d image couple%s to compute.", len(indices),
                    "" if len(indices) == 1 else "s")
        if self.work_piv is None:
            self.work_piv = self.WorkVelocimetry(self.params)
        im2im = self._im2im()
        g = self.params.gpu
        bsize = max(1, int(g.batch_pairs))
        batches = [indices[i:i + bsize] for i in range(0, len(indices), bsize)]
        params_dict = self.params._as_dict()
        lock = threading.Lock()
        with ThreadPoolExecutor(max(1, int(g.nb_io_workers))) as readers, \
                ThreadPoolExecutor(1) as saver:
            ahead = max(1, int(g.read_ahead_batches))
            pending = {}

            def submit(ib):
                if ib < len(batches) and ib not in pending:
                    pending[ib] = [readers.submit(self._read, k) for k in batches[ib]]

            for ib in range(min(ahead, len(batches))):
                submit(ib)
            saves = []
            for ib in range(len(batches)):
                submit(ib)
                futures = pending.pop(ib)
                for j in range(ib + 1, ib + 1 + ahead):
                    submit(j)
                couples = [f.result() for f in futures]
                if im2im is not None:
                    couples = [(k, im2im(a), im2im(b)) for k, a, b in couples]
                # same-shape groups of the batch, each one pipeline run
                groups = {}
                for k, a, b in couples:
                    groups.setdefault((a.shape, a.dtype.str, b.shape), []).append((k, a, b))
                for group in groups.values():
                    t0 = time.perf_counter()
                    try:
                        res = self._compute_group(group)
                    except Exception as exc:  # noqa: BLE001 (one couple alone raised)
                        (k, _, _), = group
                        logger.error("couple %s: %s: %s", self.series.names[k],
                                     type(exc).__name__, exc)
                        self.errors[self.series.names[k]] = exc
                        if stop_if_error:
                            raise
                        continue
                    self.durations["compute_piv"].append(time.perf_counter() - t0)
                    for (k, _, _), r in zip(group, res):
                        if r is not None:
                            saves.append(saver.submit(self._save, k, r, params_dict, lock))
                        elif stop_if_error:
                            raise self.errors[self.series.names[k]]
            for f in saves:
                f.result()
        dt = time.perf_counter() - t_start
        logger.info("Stop compute after t = %.2f s (%d piv fields, %.3f s/field)", dt,
                    len(self.results), dt / max(1, len(self.results)))

    def _compute_group(self, group):
        """Results of the group's couples (None for a couple whose work
        raised: the error is logged and kept in ``self.errors``, and the
        series goes on, as the reference's executors go on after a failed
        work)."""
        from ..data_objects import ArrayCouple

        pairs = []
        for k, a, b in group:
            f0, f1 = self.series.couples[k]
            c = ArrayCouple(names=(f0, f1), arrays=(a, b))
            c.apply_mask(self.params.mask)
            pairs.append(c.get_arrays())
        try:
            return self.work_piv.calcul_batch(pairs, keep_maps=False)
        except Exception:
            if len(pairs) == 1:
                raise
        out = []
        for (k, _, _), pair in zip(group, pairs):
            try:
                out.append(self.work_piv.calcul_batch([pair], keep_maps=False)[0])
            except Exception as exc:  # noqa: BLE001 (any failed work is reported)
                logger.error("couple %s: %s: %s", self.series.names[k], type(exc).__name__, exc)
                self.errors[self.series.names[k]] = exc
                out.append(None)
        return out

    def _save(self, k, results, params_dict, lock):
        t0 = time.perf_counter()
        path = self.path_dir_result / (self.series.names[k] + ".npz")
        save_results(path, results, self.series.couples[k], params_dict)
        with lock:
            self.results.append(path)
            self.durations["save_piv"].append(time.perf_counter() - t0)


Topology = TopologyPIV
