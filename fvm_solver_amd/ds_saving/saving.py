"""Trajectory saver -- the file format of ``time_fvm/ds_saving/saving.py`` (Saver,
load_step, saving.py:11-88) written without stalling the GPU step loop.

Files: ``mesh_props.npz`` (triangles, vertices, centroids, edges, bc_midpoints,
bc_normals, the boundary-edge mask under the reference's key ``bc_edge_masK``,
bc_type_str) once, then one ``t_{t:.4g}.npz`` per save with t, prim_mean, prim_std
(per-column mean and unbiased std of the cell primitives) and the fp16
``cell_primatives`` and ``bc_primatives`` normalised by them (std clamped at 1e-8).
``bc_primatives`` are the boundary facets' primitives averaged over the facet's two
sides (the interior cell's limited reconstruction and the boundary value) of the last
RHS evaluation, as the reference takes them from ``E_props.*_facet``.

The normalisation runs on the device; the fp16 arrays are copied into pinned host
buffers on a side stream and a worker thread compresses and writes the file, so a save
costs the step loop a few small launches and no host synchronisation.
"""
from __future__ import annotations

import os
import queue
import threading
from datetime import datetime

import numpy as np
import torch


class Saver:
    def __init__(self, E_props, save_dir=None):
        if save_dir is None:
            timestamp = datetime.now().strftime("%m-%d_%H-%M-%S")
            save_dir = os.path.join(os.environ.get("FVM_ARTEFACT_DIR", "artefacts"), "fvm_saves", timestamp)
        self.save_dir = save_dir
        mesh = E_props.fvm_mesh
        dm = E_props.mesh
        bcm = mesh.bc_facet_mask.cpu().numpy()
        self.mesh_props = {
            "triangles": mesh.cells.cpu().numpy(), "vertices": mesh.vertices.cpu().numpy(),
            "centroids": mesh.centroids.cpu().numpy(), "edges": mesh.facets.cpu().numpy(),
            "bc_midpoints": mesh.midpoints[torch.from_numpy(bcm)].numpy(),
            "bc_normals": mesh.normals[torch.from_numpy(bcm)].numpy(),
            "bc_edge_masK": bcm, "bc_type_str": np.asarray(E_props.bc_type_str)}
        self.saved_mesh_props = False
        # static description of each boundary facet's interior side
        dev = E_props.device
        self.D, self.C = E_props.dim, E_props.n_comp
        h = dm.host
        bcell = h["bcell"]
        fid = np.flatnonzero(bcm)
        c2f = mesh.cell_to_facet.numpy()
        k = np.argmax(c2f[bcell] == fid[:, None], axis=1)
        disp = mesh.cent_to_facet_disp.numpy()[bcell, k]
        self._bcell = torch.from_numpy(bcell.astype(np.int64)).to(dev)
        self._disp = torch.from_numpy(disp).float().to(dev)
        self._stream = torch.cuda.Stream(device=dev)
        self._q: queue.Queue = queue.Queue()
        self._worker = threading.Thread(target=self._write_loop, daemon=True)
        self._worker.start()
        self._errors: list = []

    # ------------------------------------------------------------------ device side
    def bc_values(self, E_props, prim):
        """[F_b, C] mean over the two sides of each boundary facet of the last RHS
        evaluation's facet primitives (prim = that evaluation's cell primitives)."""
        D, C = self.D, self.C
        g = E_props.gp.view(-1, D + 1, C)[self._bcell]          # [Fb, D+1, C]
        P = prim[self._bcell]
        dU = (g[:, :D, :] * self._disp[:, :, None]).sum(1)       # grad . r
        inner = P + dU * g[:, D, :]
        return 0.5 * (inner + E_props.ubc[:len(self._bcell)])

    def save(self, t, E_props, primatives):
        if not self.saved_mesh_props:
            os.makedirs(self.save_dir, exist_ok=True)
            self._q.put(("mesh_props.npz", dict(self.mesh_props)))
            self.saved_mesh_props = True
        last = getattr(E_props, "last_prim", None)
        bc_vals = self.bc_values(E_props, last if last is not None else primatives)
        prim_mean = primatives.mean(dim=0, keepdim=True)
        prim_std = primatives.std(dim=0, keepdim=True)
        safe = prim_std.clamp(min=1e-8)
        cell16 = ((primatives - prim_mean) / safe).half()
        bc16 = ((bc_vals - prim_mean) / safe).half()
        tt = t if isinstance(t, torch.Tensor) else torch.tensor(float(t), device=primatives.device)
        cur = torch.cuda.current_stream(primatives.device)
        self._stream.wait_stream(cur)
        with torch.cuda.stream(self._stream):
            host = {k: torch.empty(v.shape, dtype=v.dtype, pin_memory=True)
                    for k, v in (("t", tt.reshape(1).double()), ("prim_mean", prim_mean), ("prim_std", prim_std),
                                 ("cell_primatives", cell16), ("bc_primatives", bc16))}
            srcs = {"t": tt.reshape(1).double(), "prim_mean": prim_mean, "prim_std": prim_std,
                    "cell_primatives": cell16, "bc_primatives": bc16}
            for k, v in srcs.items():
                host[k].copy_(v, non_blocking=True)
                v.record_stream(self._stream)
            ev = torch.cuda.Event()
            ev.record(self._stream)
        self._q.put((ev, host))

    # ------------------------------------------------------------------ host side
    def _write_loop(self):
        while True:
            item = self._q.get()
            if item is None:
                return
            try:
                if isinstance(item[0], str):
                    np.savez_compressed(os.path.join(self.save_dir, item[0]), **item[1])
                else:
                    ev, host = item
                    ev.synchronize()
                    t = float(host["t"][0])
                    values = {"t": t, "prim_mean": host["prim_mean"].numpy(), "prim_std": host["prim_std"].numpy(),
                              "cell_primatives": host["cell_primatives"].numpy(),
                              "bc_primatives": host["bc_primatives"].numpy()}
                    np.savez_compressed(os.path.join(self.save_dir, f"t_{t:.4g}.npz"), **values)
            except Exception as e:  # reported by flush()
                self._errors.append(e)
            finally:
                self._q.task_done()

    def flush(self):
        """Wait until every queued file is written (raises the first write error)."""
        self._q.join()
        if self._errors:
            raise self._errors[0]

    def close(self):
        self.flush()
        self._q.put(None)
        self._worker.join()


def load_step(file_path):
    """Un-normalised primitives of one saved step: (t, cell [N, C], bc [F_b, C])."""
    data = np.load(file_path, allow_pickle=False)
    mean, std = data["prim_mean"], data["prim_std"]
    cell = data["cell_primatives"].astype(np.float32) * std + mean
    bc = data["bc_primatives"].astype(np.float32) * std + mean
    return data["t"], torch.from_numpy(cell).float(), torch.from_numpy(bc).float()
