"""Device-resident multipass PIV over a batch of image pairs.

One pass = one launch of ``piv_make_windows``, ONE launch of the fused
correlation kernel over the windows of every pair, one ``piv_fix``, and the interpolation of the fixed field on the next grid
(GPU FP64 TPS, or scipy's griddata exactly as the reference when
``use_tps`` selects it). The semantics are those of
src/fluidimage/works/piv/multipass.py:172-190 applied to each pair; results
stay on the device until ``to_host``.
"""

import numpy as np

from ... import _native
from ...data_objects import HeavyPIVResults, MultipassPIVResults
from ...device import ImageSet, to_device_image
from .fix import run_fix_device


class PassOutput:
    """Device tensors of one pass for a batch (pair-major slices)."""

    def __init__(self, work, npairs, out, win, dev):
        self.work = work
        self.npairs = npairs
        self.nvec = work.nb_vectors
        self.out = out
        self.win = win
        self.dev = dev
        self.approx = [None] * npairs  # predictor on the next grid, per pair
        self.final = [None] * npairs
        self.refined = [None] * npairs
        self.tps = [None] * npairs

    def pair(self, key, k):
        t = self.out[key] if key in self.out else self.win[key]
        return t[k * self.nvec:(k + 1) * self.nvec]


class MultipassPipeline:
    """Runs the passes of a ``WorkPIV`` on batches of same-shape pairs."""

    def __init__(self, work_piv):
        self.work = work_piv
        self.params = work_piv.params
        self._deformed = {}
        self._refine_problem = {}
        self._refine_fields = {}
        # the TPS calls do not wait for their singular-system check: it runs
        # once at the end of ``run``, and a pair that failed it is run again
        # alone with the checks in place (griddata fallback and all)
        self.defer_checks = True
        self._pending = []
        # "device": TPS batches sized on the GPU from capacities planned once
        # per pass (no host read inside a run; a field beyond them is run
        # again with "host" sizing); "host": sized from device-to-host reads
        self.sizing = "device"
        self._caps = {}
        self._tables = {}
        self._grids = {}
        # device-sized TPS on grid-node centers: subdomains with one node
        # pattern share a factorisation (a capacitance correction for their
        # missing nodes); False: every subdomain factorised alone
        self.share_factors = True

    def _event_pair(self, timers, key):
        import torch

        if timers is None:
            return None
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        timers.setdefault(key, []).append(ev)
        return ev

    def _interp(self, work_next, res, last, stream=None, timers=None):
        """apply_interp (singlepass.py:409-563) of every pair's fixed field
        onto work_next's grid, returning per pair the float64 device tensors
        (x, y). The TPS of all pairs runs as one batch; a pair whose TPS
        system is singular falls back to griddata, as in the reference."""
        from ...calcul.tps import SingularSystemError, lattice_step, tps_device_batch

        mp = self.params.multipass
        use_tps = mp.use_tps is True or (mp.use_tps == "last" and last)
        out = [None] * res.npairs
        if use_tps and self.sizing == "device":
            return self._interp_planned(work_next, res, last, stream, timers)
        if use_tps:
            dev = res.dev
            gx, gy = work_next._device_grid_full(dev)
            fields = [(res.pair("dx", k), res.pair("dy", k), res.pair("xs", k), res.pair("ys", k))
                      for k in range(res.npairs)]
            ev = self._event_pair(timers, f"tps_pass{work_next.index_pass}" + ("_final" if last else ""))
            if ev:
                ev[0].record()
            # the centers of a pass sit on its grid (deformed passes; the first
            # pass) and the new positions are the next grid: the common step of
            # both grids is the lattice of the kernel-function table (positions
            # off that lattice are evaluated directly)
            step = lattice_step(res.work.ixvecs, res.work.iyvecs, work_next.ixvecs,
                                work_next.iyvecs)
            tps = tps_device_batch(fields, gx, gy, mp.subdom_size, mp.smoothing_coef,
                                   mp.threshold_tps, 20, stream=stream, lattice=step,
                                   defer=self.defer_checks)
            if ev:
                ev[1].record()
            for k, t in enumerate(tps):
                if isinstance(t, SingularSystemError):
                    if "Only nan" not in str(t):
                        print("LinAlgError in TPS: compute delta_approx with griddata")
                    continue  # griddata below (which raises "Only nan." itself)
                res.tps[k] = t
                out[k] = (t["approx"][0], t["approx"][1])
                if "pending" in t:
                    self._pending.append((k, t))
        for k in range(res.npairs):
            if out[k] is None:
                out[k] = self._griddata(work_next, res, k)
        return out

    def _capacity(self, work, work_next, npairs, deformed):
        """Capacities of the device-sized TPS call from ``work``'s vectors to
        ``work_next``'s grid (planned once per pair of grids and batch size)."""
        from ...calcul.tps import lattice_step
        from ...calcul.tps_planned import make_lattice_table, plan_capacity

        key = (work.index_pass, work_next.index_pass, npairs, deformed, work.imshape0)
        cap = self._caps.get(key)
        if cap is None:
            mp = self.params.multipass
            step = lattice_step(work.ixvecs, work.iyvecs, work_next.ixvecs, work_next.iyvecs)
            # the first pass and the deformed passes put their vectors on the grid
            on_grid = deformed or work.index_pass == 0
            cap = plan_capacity(work.ixvecs_grid, work.iyvecs_grid, npairs, mp.subdom_size,
                                mp.smoothing_coef, 20, step=step, generated=on_grid)
            self._caps[key] = cap
        if cap not in self._tables:
            import torch

            self._tables[cap] = make_lattice_table(cap, torch.device("cuda"))
        return cap, self._tables[cap]

    def _pass_grid(self, work, dev):
        from ...calcul.tps_planned import PassGrid

        key = (work.index_pass, work.imshape0)
        g = self._grids.get(key)
        if g is None:
            g = self._grids[key] = PassGrid(work.ixvecs, work.iyvecs, dev)
        return g

    def _share(self, work, cap, deformed):
        """Whether the TPS call from ``work``'s vectors shares factorisations:
        only when the vectors sit on the pass grid (the first pass, deformed
        passes) and the batch has many subdomains, many per field (a pattern
        then serves many of them)."""
        from ...calcul import tps_planned as tp

        return (bool(self.share_factors) and (deformed or work.index_pass == 0)
                and cap.dom_cap >= tp.SHARED_MIN_SUBDOMAINS and cap.nd >= tp.SHARED_MIN_BATCH)

    def _interp_planned(self, work_next, res, last, stream=None, timers=None):
        """The TPS interpolation of every pair as one device-sized batch (no
        host read); the fields' checks are read at the end of the run."""
        from ...calcul.tps_planned import tps_planned

        mp = self.params.multipass
        deformed = getattr(self.params.multipass, "image_deformation", None) is not None \
            and res.work.index_pass > 0
        cap, table = self._capacity(res.work, work_next, res.npairs, deformed)
        gx, gy = work_next._device_grid_full(res.dev)
        ev = self._event_pair(timers, f"tps_pass{work_next.index_pass}" + ("_final" if last else ""))
        if ev:
            ev[0].record()
        grid = None
        if self._share(res.work, cap, deformed):
            # the vectors sit on the pass grid: subdomains share factorisations
            grid = self._pass_grid(res.work, res.dev)
        t = tps_planned(cap, res.out["dx"], res.out["dy"], res.win["xs"], res.win["ys"], gx, gy,
                        mp.subdom_size, mp.smoothing_coef, mp.threshold_tps, 20, stream=stream,
                        lattice_table=table, grid=grid)
        if ev:
            ev[1].record()
        self._pending.append(("planned", t))
        out = []
        for k in range(res.npairs):
            res.tps[k] = (t, k)
            out.append(t.approx(k))
        return out

    def _griddata(self, work_next, res, k):
        import torch

        from .interp import griddata

        dx = res.pair("dx", k)
        dy = res.pair("dy", k)
        dev = dx.device
        hdx = dx.cpu().numpy()
        hdy = dy.cpu().numpy()
        sel = ~(np.isnan(hdx) | np.isnan(hdy))
        if not sel.any():
            raise ValueError("Only nan.")
        hx = res.pair("xs", k).cpu().numpy()[sel]
        hy = res.pair("ys", k).cpu().numpy()[sel]
        centers = np.vstack([hy, hx])
        ax = griddata(centers[::-1], hdx[sel], (work_next.ixvecs, work_next.iyvecs))
        ay = griddata(centers[::-1], hdy[sel], (work_next.ixvecs, work_next.iyvecs))
        return (torch.from_numpy(np.ascontiguousarray(ax)).to(dev),
                torch.from_numpy(np.ascontiguousarray(ay)).to(dev))

    def run(self, imageset, npairs, keep_maps=False, stream=None, final_interp=True,
            timers=None, check=True):
        """All passes for pairs (2k, 2k+1) of ``imageset``; returns the list
        of PassOutput. ``timers`` (dict) collects event pairs around the
        correlation launches and the TPS calls. With a ``stream``, all work
        (launches, allocations, copies, host reads) is ordered on it.
        ``check=False`` leaves the TPS checks of the run in ``self._pending``
        (the run then reads nothing back from the device): the caller checks
        them later with ``check_pending``."""
        from ...device import on_stream

        with on_stream(stream):
            return self._run(imageset, npairs, keep_maps, stream, final_interp, timers, check)

    def _run(self, imageset, npairs, keep_maps, stream, final_interp, timers, check=True):
        import torch

        from .deform import DeformedPass, add_predictor, check_method

        works = self.work.works_piv
        dev = imageset.device
        passes = []
        approx = [None] * npairs
        self._pending = []
        deform = check_method(getattr(self.params.multipass, "image_deformation", None))
        if deform and self.params.piv0.nb_peaks_to_search > 1:
            raise NotImplementedError(
                "image deformation with nb_peaks_to_search > 1 is not implemented")
        for ip, work in enumerate(works):
            nvec = work.nb_vectors
            deformed = deform is not None and ip > 0
            table, dtype = imageset.table, imageset.dtype
            if deformed:
                # windows centred on the grid nodes, resampled by the predictor
                u = torch.cat([approx[k][0] for k in range(npairs)])
                v = torch.cat([approx[k][1] for k in range(npairs)])
                dp = self._deformed.get(ip)
                if dp is None:
                    dp = self._deformed[ip] = DeformedPass(work, dev)
                ev = self._event_pair(timers, f"deform_pass{ip}")
                if ev:
                    ev[0].record()
                deformed_images = dp.run(imageset, npairs, u, v, stream=stream)
                if ev:
                    ev[1].record()
                table, dtype = deformed_images.table, deformed_images.dtype
                windows = work.make_windows(0, 1, None, None, device=dev, stream=stream,
                                            npairs=npairs)["windows"]
                gx, gy = work._device_grid_full(dev)
                win = {"xs": gx.repeat(npairs), "ys": gy.repeat(npairs), "dx_in": u, "dy_in": v}
            else:
                dxa = dya = None
                if approx[0] is not None:
                    dxa = torch.cat([approx[k][0] for k in range(npairs)])
                    dya = torch.cat([approx[k][1] for k in range(npairs)])
                w = work.make_windows(0, 1, dxa, dya, device=dev, stream=stream, npairs=npairs)
                windows = w["windows"]
                win = {key: w[key] for key in ("xs", "ys", "dx_in", "dy_in")}
            maps = None
            if keep_maps:
                maps = torch.empty((npairs * nvec,) + work.shape_crop_im0, dtype=torch.float32,
                                   device=dev)
            ev = self._event_pair(timers, f"correlate_pass{ip}")
            if ev:
                ev[0].record()
            out = work.correl.correlate_windows(table, windows, npairs * nvec, dtype,
                                                work.has_to_apply_subpix, maps=maps,
                                                stream=stream)
            if ev:
                ev[1].record()
            if deformed:
                add_predictor(out["dx"], out["dy"], win["dx_in"], win["dy_in"], stream=stream)
            out["maps"] = maps
            out["windows"] = windows
            res = PassOutput(work, npairs, out, win, dev)
            ny, nx = len(work.iyvecs), len(work.ixvecs)
            run_fix_device(out, ny, nx, self.params.fix, stream=stream, npairs=npairs)
            if self.params.piv0.nb_peaks_to_search >= 2:
                self._secondary_peaks(res)
            passes.append(res)
            if ip + 1 < len(works):
                approx = self._interp(works[ip + 1], res, last=False, stream=stream,
                                      timers=timers)
                for k in range(npairs):
                    res.approx[k] = approx[k]
        if final_interp:
            res = passes[-1]
            try:
                final = self._interp(works[-1], res, last=True, stream=stream, timers=timers)
            except ValueError as e:
                print("Warning: InterpError at the end of the last piv pass:", e)
            else:
                for k in range(npairs):
                    res.final[k] = final[k]
                steps = int(getattr(self.params.multipass, "refine_steps", 0) or 0)
                if steps > 0:
                    self._refine(imageset, res, steps, timers)
        if check:
            failed = self._check_deferred()
            if failed:
                self._rerun_pairs(imageset, passes, failed, keep_maps, stream, final_interp)
        return passes

    def _check_deferred(self):
        """Pairs whose deferred TPS check failed (one device-to-host read per
        TPS call of the run); clears the pending checks."""
        failed = self.check_pending(self._pending)
        self._pending = []
        return failed

    @staticmethod
    def check_pending(pending):
        """Pairs failing any of the given deferred TPS checks (a singular
        system, a field beyond the device-sized capacities, ...)."""
        from ...calcul.tps import deferred_error

        failed = set()
        for k, t in pending:
            if k == "planned":
                failed |= {j for j, e in enumerate(t.field_errors()) if e is not None}
            elif deferred_error(t) is not None:
                failed.add(k)
        return failed

    def _rerun_pairs(self, imageset, passes, pairs, keep_maps, stream, final_interp):
        """Run the given pairs again one by one with the TPS checks in place
        and put their results into the batch's passes (a pair's results do
        not depend on the other pairs of its batch)."""
        import torch

        for k in sorted(pairs):
            sub = MultipassPipeline(self.work)
            sub.defer_checks = False
            sub.sizing = "host"
            single = ImageSet([imageset.images[2 * k], imageset.images[2 * k + 1]],
                              stream=stream)
            one_passes = sub.run(single, 1, keep_maps=keep_maps, stream=stream,
                                 final_interp=final_interp)
            for res, one in zip(passes, one_passes):
                n = res.nvec
                for store, src in ((res.out, one.out), (res.win, one.win)):
                    for key, t in store.items():
                        s1 = src.get(key)
                        if (key == "windows" or not torch.is_tensor(t) or s1 is None
                                or t.dim() == 0 or t.shape[0] != res.npairs * n):
                            continue
                        t[k * n:(k + 1) * n].copy_(s1)
                res.approx[k] = one.approx[0]
                res.final[k] = one.final[0]
                res.refined[k] = one.refined[0]
                res.tps[k] = one.tps[0]
                if getattr(one, "replaced", None):
                    res.replaced = getattr(res, "replaced", {})
                    res.replaced[k] = one.replaced.get(0, {})

    def _refine(self, imageset, res, steps, timers=None):
        """Differentiable refinement (refine.py) of the final fields of all
        pairs from the multipass result (the fallback: a step is kept only
        when it lowers its pair's loss)."""
        import torch

        from .refine import RefineProblem, node_field, refine

        key = id(imageset)
        prob = self._refine_problem.get(key)
        if prob is None or prob.npairs != res.npairs:
            prob = RefineProblem(res.work, imageset, res.npairs)
            self._refine_problem = {key: prob}
        mp = self.params.multipass
        field = None
        if getattr(mp, "refine_field", "tps") == "tps":
            fkey = (res.work.index_pass, res.work.imshape0, mp.subdom_size, mp.smoothing_coef)
            field = self._refine_fields.get(fkey)
            if field is None:
                field = self._refine_fields[fkey] = node_field(prob, mp.subdom_size,
                                                               mp.smoothing_coef)
        ev = self._event_pair(timers, "refine")
        if ev:
            ev[0].record()
        dx0 = torch.stack([res.final[k][0] for k in range(res.npairs)])
        dy0 = torch.stack([res.final[k][1] for k in range(res.npairs)])
        dx, dy, l0, l1, acc = refine(prob, dx0, dy0, steps=steps, field=field)
        if ev:
            ev[1].record()
        for k in range(res.npairs):
            res.refined[k] = (dx[k], dy[k], float(l0[k]), float(l1[k]), int(acc[k]))

    def _secondary_peaks(self, res):
        """WorkFIX secondary-peak replacement (fix.py:135-243) on the host,
        written back to the device fields."""
        import torch

        from .fix import WorkFIX

        for k in range(res.npairs):
            r = self._pass_to_heavy(res, k, fixed=True, with_maps=False)
            h = {key: res.pair(key, k).cpu().numpy()
                 for key in ("dx", "dy", "flags", "diff2", "mean_x", "mean_y")}
            fixer = WorkFIX(self.params.fix, res.work)
            fixer._secondary(r, h, {}, {})
            replaced = getattr(r, "replaced_vectors", {})
            if not replaced:
                continue
            res.pair("dx", k).copy_(torch.from_numpy(np.asarray(r.deltaxs, np.float32)))
            res.pair("dy", k).copy_(torch.from_numpy(np.asarray(r.deltays, np.float32)))
            res.pair("correl_max", k).copy_(
                torch.from_numpy(np.asarray(r.correls_max, np.float32)))
            # a replaced vector loses its error entry (fix.py:231-236)
            cleared = [i for i in replaced if i not in r.errors]
            if cleared:
                ids = torch.tensor(cleared, dtype=torch.long, device=res.dev)
                res.pair("flags", k)[ids] = 0
                res.pair("err", k)[ids] = 0
            res.replaced = getattr(res, "replaced", {})
            res.replaced[k] = replaced

    # ------------------------------------------------------------------ host
    def _pass_to_heavy(self, res, k, fixed=True, with_maps=True, couple=None):
        work = res.work
        g = lambda key: res.pair(key, k).cpu().numpy()  # noqa: E731
        err = g("err")
        flags = g("flags") if fixed else None
        from ...data_objects import error_strings

        errors = error_strings(err, flags, g("diff2") if fixed else None)
        if res.work.index_pass == 0:
            xs = work.ixvecs_grid.copy()
            ys = work.iyvecs_grid.copy()
        else:
            xs = g("xs")
            ys = g("ys")
        xs, ys = work._xyoriginalimage_from_xymasked(xs, ys)
        correls = None
        if with_maps and res.out.get("maps") is not None:
            correls = list(res.pair("maps", k).cpu().numpy())
        nb = self.params.piv0.nb_peaks_to_search
        secondary = [None] * len(err)
        sec_sub = None
        if nb > 1:
            sec = g("sec")
            nsec = g("nsec")
            sec_sub = g("sec_sub")
            for i in range(len(err)):
                if err[i] in (_native.WIN_OK, _native.WIN_SUBPIX_WRONG,
                              _native.WIN_SUBPIX_BOUNDARY):
                    secondary[i] = [(int(a), int(b), np.float32(c))
                                    for (a, b, c) in sec[i, :nsec[i]]]
        r = HeavyPIVResults(g("dx"), g("dy"), xs, ys, errors, correls_max=g("correl_max"),
                            correls=correls, couple=couple, params=self.params,
                            secondary_peaks=secondary)
        r._sec_sub = sec_sub
        r.indices_no_displacement = work.correl.get_indices_no_displacement()
        r.displacement_max = work.correl.displacement_max
        if work.index_pass > 0:
            r.deltaxs_input = g("dx_in")
            r.deltays_input = g("dy_in")
        return r

    def to_host(self, passes, k, couple=None):
        """MultipassPIVResults of pair k (reference attribute names)."""
        results = MultipassPIVResults()
        works = self.work.works_piv
        for ip, res in enumerate(passes):
            r = self._pass_to_heavy(res, k, couple=couple)
            if getattr(res, "replaced", None):
                r.replaced_vectors = res.replaced.get(k, {})
            work = res.work
            if ip + 1 < len(passes) and res.approx[k] is not None:
                nxt = works[ip + 1]
                r.ixvecs_approx, r.iyvecs_approx = nxt._xyoriginalimage_from_xymasked(
                    nxt.ixvecs_grid, nxt.iyvecs_grid)
                r.deltaxs_approx = res.approx[k][0].cpu().numpy()
                r.deltays_approx = res.approx[k][1].cpu().numpy()
            if ip + 1 == len(passes) and res.final[k] is not None:
                r.ixvecs_final, r.iyvecs_final = work._xyoriginalimage_from_xymasked(
                    work.ixvecs_grid, work.iyvecs_grid)
                r.deltaxs_final = res.final[k][0].cpu().numpy()
                r.deltays_final = res.final[k][1].cpu().numpy()
                t = res.tps[k]
                if t is not None:
                    if isinstance(t, tuple):  # (device-sized batch, field)
                        sel, sm = t[0].host_field(t[1])
                    else:
                        sel = t["index"].cpu().numpy()
                        sm = t["smoothed"].cpu().numpy()
                    xs_s = r.ixvecs_final.copy().astype(np.float32)
                    ys_s = r.iyvecs_final.copy().astype(np.float32)
                    xs_s[sel] = r.xs[sel]
                    ys_s[sel] = r.ys[sel]
                    fx = r.deltaxs_final.copy()
                    fy = r.deltays_final.copy()
                    fx[sel] = sm[0]
                    fy[sel] = sm[1]
                    r.xs_smooth, r.ys_smooth = xs_s, ys_s
                    r.deltaxs_smooth, r.deltays_smooth = fx, fy
                if res.refined[k] is not None:
                    rx, ry, l0, l1, acc = res.refined[k]
                    r.deltaxs_refined = rx.cpu().numpy()
                    r.deltays_refined = ry.cpu().numpy()
                    r.refine_loss = (l0, l1)
                    r.refine_accepted_steps = acc
            results.append(r)
        return results


class GraphedPasses:
    """Every pass of a ``WorkPIV`` over a batch of pairs (image minima,
    window records, the fused correlation of each pass, WorkFIX, the
    device-sized TPS interpolation between passes and at the end, the image
    deformation) captured once in a HIP graph and replayed: a replay re-runs
    every kernel on the images currently in ``imageset`` (same tensors, new
    contents) and writes into the same result tensors (``passes``).

    Nothing in the captured step reads the device: the TPS checks (singular
    systems, fields beyond the capacities planned for the batch) are read
    after each replay, once, and a pair that failed one is run again alone
    with host sizing (the reference's fallbacks), its results copied into
    the replay's. Needs TPS for every interpolation (use_tps True; griddata
    runs on the host), one peak per window and no refinement steps."""

    def __init__(self, work_piv, imageset, npairs, keep_maps=False, warmup=2,
                 final_interp=None):
        import torch

        p = work_piv.params
        mp = p.multipass
        nwork = len(work_piv.works_piv)
        if final_interp is None:
            final_interp = nwork > 1
        needs_tps = nwork > 1 or final_interp
        use_tps_all = mp.use_tps is True or (mp.use_tps == "last" and nwork == 1)
        if needs_tps and not use_tps_all:
            raise NotImplementedError(
                "graph capture of the multipass: griddata (use_tps False or 'last' between "
                "passes) runs on the host; use params.multipass.use_tps = True")
        if p.piv0.nb_peaks_to_search > 1:
            raise NotImplementedError("secondary-peak replacement runs on the host")
        if int(getattr(mp, "refine_steps", 0) or 0) > 0:
            raise NotImplementedError("the refinement steps accept or reject on the host")
        self.pipe = MultipassPipeline(work_piv)
        self.pipe.sizing = "device"
        self.imageset = imageset
        self.npairs = npairs
        self.keep_maps = keep_maps
        self.final_interp = final_interp
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):  # plans, grids, capacities and caches outside the graph
            for _ in range(warmup):
                self._step(check=True)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.passes = self._step(check=False)
        self.pending = list(self.pipe._pending)
        self.pipe._pending = []
        # the graph's own output references (a rerun replaces some of them)
        self._refs = [(list(r.approx), list(r.final), list(r.tps)) for r in self.passes]
        self.reruns = 0

    def _step(self, check):
        self.imageset.recompute_min()
        return self.pipe.run(self.imageset, self.npairs, keep_maps=self.keep_maps,
                             final_interp=self.final_interp, check=check)

    def replay(self):
        """Run the captured step on the current stream, then read its TPS
        checks (one device-to-host read) and run the pairs that failed one
        again. Returns the passes holding the results."""
        self.graph.replay()
        for res, (approx, final, tps) in zip(self.passes, self._refs):
            res.approx[:], res.final[:], res.tps[:] = approx, final, tps
        failed = self.pipe.check_pending(self.pending)
        if failed:
            self.reruns += len(failed)
            self.pipe._rerun_pairs(self.imageset, self.passes, failed, self.keep_maps, None,
                                   self.final_interp)
        return self.passes


class UnevenGraph:
    """The uneven batch of ``run_uneven`` (pairs of different shapes and
    window schedules) as one captured HIP graph: each (shape, schedule)
    bucket is a batched multipass, and the buckets are captured one after
    the other into the same graph; ``replay`` runs them all. ``items``:
    (im0, im1, schedule) with the images already on the device (their
    contents may change between replays). With ``concurrent`` the buckets
    are captured on streams of their own, forked from and joined to the
    capture stream, so the graph holds them as parallel branches: the
    GPU runs the small buckets' kernels side by side instead of one bucket
    after the other (each bucket's own kernels keep their order)."""

    def __init__(self, items, params, keep_maps=False, warmup=2, concurrent=True, split=1):
        from copy import deepcopy

        import torch

        from ...device import ImageSet, to_device_image
        from .multipass import WorkPIV
        from .schedule import schedule_params

        dev = torch.device("cuda")
        groups = {}
        for i, (im0, _im1, sched) in enumerate(items):
            groups.setdefault((tuple(im0.shape), tuple(sched)), []).append(i)
        # ``split``: each bucket as that many sub-batches (parallel branches
        # with ``concurrent``: one sub-batch's latency-bound kernels then run
        # beside another's compute-bound ones)
        chunks = []
        for key, idx in groups.items():
            k = max(1, min(int(split), len(idx)))
            bounds = [len(idx) * j // k for j in range(k + 1)]
            chunks += [(key, idx[bounds[j]:bounds[j + 1]]) for j in range(k)]
        self.groups = []
        for (shape, sched), idx in chunks:
            work = WorkPIV(schedule_params(deepcopy(params), sched))
            for w in work.works_piv:
                w._prepare_with_image(imshape=shape)
            ims = []
            for i in idx:
                ims += [to_device_image(items[i][0], dev), to_device_image(items[i][1], dev)]
            self.groups.append((work, ImageSet(ims), idx))
        self.pipes = []
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for work, images, idx in self.groups:
                pipe = MultipassPipeline(work)
                pipe.sizing = "device"
                for _ in range(warmup):
                    images.recompute_min()
                    pipe.run(images, len(idx), keep_maps=keep_maps)
                self.pipes.append(pipe)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.keep_maps = keep_maps
        self.graph = torch.cuda.CUDAGraph()
        self.passes = []
        # default-priority streams, one per bucket
        self.streams = [torch.cuda.Stream() for _ in self.groups] if concurrent else []
        with torch.cuda.graph(self.graph):
            cap = torch.cuda.current_stream()
            for g, ((work, images, idx), pipe) in enumerate(zip(self.groups, self.pipes)):
                branch = self.streams[g] if concurrent else cap
                branch.wait_stream(cap)
                with torch.cuda.stream(branch):
                    images.recompute_min()
                    self.passes.append(pipe.run(images, len(idx), keep_maps=keep_maps,
                                                check=False))
            for branch in self.streams:
                cap.wait_stream(branch)
        self.pending = []
        for pipe in self.pipes:
            self.pending.append(list(pipe._pending))
            pipe._pending = []
        self._refs = [[(list(r.approx), list(r.final), list(r.tps)) for r in passes]
                      for passes in self.passes]
        self.nitems = len(items)

    def replay(self):
        """Run every bucket of the captured batch; pairs failing a TPS check
        are run again alone. Returns the passes of each bucket."""
        self.graph.replay()
        for g, ((work, images, idx), pipe) in enumerate(zip(self.groups, self.pipes)):
            for res, (approx, final, tps) in zip(self.passes[g], self._refs[g]):
                res.approx[:], res.final[:], res.tps[:] = approx, final, tps
            failed = pipe.check_pending(self.pending[g])
            if failed:
                pipe._rerun_pairs(images, self.passes[g], failed, self.keep_maps, None, True)
        return self.passes

    def results(self):
        """MultipassPIVResults of every pair, in the order of the items."""
        out = [None] * self.nitems
        for (work, images, idx), pipe, passes in zip(self.groups, self.pipes, self.passes):
            for j, i in enumerate(idx):
                out[i] = pipe.to_host(passes, j)
        return out


def run_pairs(work_piv, pairs, keep_maps=False, stream=None):
    """WorkPIV on a list of (im0, im1) host or device arrays of one shape."""
    import torch

    dev = torch.device("cuda")
    imgs = []
    for im0, im1 in pairs:
        imgs += [to_device_image(im0, dev), to_device_image(im1, dev)]
    shapes = {tuple(t.shape) for t in imgs}
    if len(shapes) != 1:
        raise ValueError("all images of a batch must have one shape")
    shape = shapes.pop()
    for w in work_piv.works_piv:
        if getattr(w, "imshape0", None) != shape:
            w._prepare_with_image(imshape=shape)
    from ...device import on_stream

    with on_stream(stream):
        imageset = ImageSet(imgs, stream=stream)
        pipe = MultipassPipeline(work_piv)
        passes = pipe.run(imageset, len(pairs), keep_maps=keep_maps, stream=stream)
    return pipe, passes


def run_uneven(items, params):
    """Pairs of different shapes and window schedules (a list of
    (im0, im1, schedule)): pairs sharing a shape and a schedule run as one
    batch (one correlation launch per pass, one TPS launch per kernel).
    ``params`` is a WorkPIV params tree whose piv0.shape_crop_im0 and
    multipass.number / coeff_zoom are replaced by each schedule's.
    Returns the MultipassPIVResults in the order of ``items``."""
    from copy import deepcopy

    from .multipass import WorkPIV
    from .schedule import schedule_params

    groups = {}
    for i, (im0, im1, sched) in enumerate(items):
        groups.setdefault((tuple(im0.shape), tuple(sched)), []).append(i)
    out = [None] * len(items)
    for (shape, sched), idx in groups.items():
        work = WorkPIV(schedule_params(deepcopy(params), sched))
        res = work.calcul_batch([(items[i][0], items[i][1]) for i in idx])
        for i, r in zip(idx, res):
            out[i] = r
    return out
