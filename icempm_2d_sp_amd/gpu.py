"""Host-side mirror of the reference's GPU_Implementation5 / GPU_Partition /
HostSideSOA call surface (simulation/gpu_implementation5.h:23-68,
gpu_partition.h:63-134, hssoa/host_side_soa.h:51-74) over the C-ABI.

Same method names, argument meaning and error behaviour (failures raise, as
the reference throws std::runtime_error).  One partition on one GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .params import SimParams


class HostSideSOA:
    """Host SoA [nPtsArrays][capacity] of float32 (host_side_soa.h:51-74)."""

    def __init__(self, nlines: int = abi.N_LINES):
        self.nlines = nlines
        self.host_buffer = np.zeros((nlines, 0), np.float32)
        self.capacity = 0
        self.size = 0

    def Allocate(self, pts_capacity: int):
        self.capacity = pts_capacity
        self.size = 0
        self.host_buffer = np.zeros((self.nlines, pts_capacity), np.float32)

    def getPointerToLine(self, idx_line: int) -> np.ndarray:
        return self.host_buffer[idx_line]

    def line_u32(self, idx_line: int) -> np.ndarray:
        return self.host_buffer[idx_line].view(np.uint32)

    def RemoveDisabledAndSort(self, GridY: int):
        """host_side_soa.cpp:36-47, with a stable order."""
        n = self.size
        util = self.line_u32(abi.IDX_UTILITY)[:n]
        keep = np.nonzero((util & abi.STATUS_DISABLED) == 0)[0]
        cell = self.line_u32(abi.IDX_CELL)[keep]
        key = (cell & 0xFFFF).astype(np.int64) * GridY + (cell >> 16)
        order = keep[np.argsort(key, kind="stable")]
        self.host_buffer[:, : order.size] = self.host_buffer[:, order]
        self.size = order.size

    def convertToIntegerCellFormat(self, h: float):
        """proxypoint2d.cpp:125-139 for every point: world position -> (cell, offset)."""
        n = self.size
        x = self.host_buffer[abi.IDX_POSX, :n].astype(np.float32)
        y = self.host_buffer[abi.IDX_POSX + 1, :n].astype(np.float32)
        hinv = np.float32(1.0) / np.float32(h)
        xi = (x * hinv + np.float32(0.5)).astype(np.uint32)
        yi = (y * hinv + np.float32(0.5)).astype(np.uint32)
        self.line_u32(abi.IDX_CELL)[:n] = (yi << 16) | xi
        self.host_buffer[abi.IDX_POSX, :n] = x * hinv - xi.astype(np.float32)
        self.host_buffer[abi.IDX_POSX + 1, :n] = y * hinv - yi.astype(np.float32)


class GPU_Partition:
    """One partition = one IceHandle (gpu_partition.h:63-134)."""

    prms: SimParams | None = None

    def __init__(self):
        self.lib = abi.load()
        self.h = C.c_void_p()
        self.PartitionID = 0
        self.Device = 0
        self.nPts_partition = 0
        self.GridX_partition = 0
        self.GridX_offset = 0
        self.error_code = 0
        self.disabled_points_count = 0
        self.nPtsPitch = 0
        self.nGridPitch = 0
        self.timing_10_P2GAndHalo = 0.0
        self.timing_20_acceptHalo = 0.0
        self.timing_30_updateGrid = 0.0
        self.timing_40_G2P = 0.0
        self.timing_60_ptsSent = 0.0
        self.timing_70_ptsAccepted = 0.0
        self.timing_stepTotal = 0.0
        self._initialized = False

    def _chk(self, rc, what):
        abi.check(rc, self.h, what)

    def initialize(self, device: int, partition: int):
        if self._initialized:
            raise RuntimeError("GPU_Partition double initialization")
        self.PartitionID, self.Device = partition, device
        cp = self.prms.to_c()
        rc = self.lib.ice_create(C.byref(cp), device, C.byref(self.h))
        abi.check(rc, None, "GPU_Partition initialization failure")
        self.lib.ice_enable_timings(self.h, 1)
        self._initialized = True

    def destroy(self):
        if self.h:
            self.lib.ice_destroy(self.h)
            self.h = C.c_void_p()
            self._initialized = False

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def allocate(self, n_points_capacity: int, grid_x_capacity: int):
        self._chk(self.lib.ice_allocate(self.h, n_points_capacity), "GPU_Partition allocate")
        info = self.info()
        self.nPtsPitch, self.nGridPitch = int(info.point_pitch), int(info.grid_pitch)

    def info(self) -> abi.IceInfo:
        i = abi.IceInfo()
        self._chk(self.lib.ice_get_info(self.h, C.byref(i)), "info")
        return i

    def transfer_points_from_soa_to_device(self, hssoa: HostSideSOA, point_idx_offset: int):
        buf = hssoa.host_buffer
        view = np.ascontiguousarray(buf[:, point_idx_offset: point_idx_offset + self.nPts_partition])
        self._chk(self.lib.ice_upload_points(self.h, abi.ptr(view), view.shape[1], self.nPts_partition,
                                             view.shape[0]), "transfer_points_from_soa_to_device")
        self.disabled_points_count = 0

    def transfer_grid_data_to_device(self, gpu: "GPU_Implementation5"):
        st = np.ascontiguousarray(gpu.grid_status_buffer, np.uint8)
        nrm = np.ascontiguousarray(gpu.grid_boundary_normals, np.float32) if gpu.grid_boundary_normals.size else None
        self._chk(self.lib.ice_upload_grid(self.h, abi.ptr(st), abi.ptr(nrm)), "transfer_grid_data_to_device")

    def update_constants(self):
        cp = self.prms.to_c()
        self._chk(self.lib.ice_update_constants(self.h, C.byref(cp)), "cuda_update_constants")
        self.error_code = 0

    def update_current_field(self, current: np.ndarray | None, wind_stress: np.ndarray | None = None):
        cur = None if current is None else np.ascontiguousarray(current, np.float32)
        wnd = None if wind_stress is None else np.ascontiguousarray(wind_stress, np.float32)
        self._chk(self.lib.ice_upload_forcing(self.h, abi.ptr(cur), abi.ptr(wnd)), "update_current_field")

    def transfer_from_device(self, hssoa: HostSideSOA, point_idx_offset: int, boundary_forces: np.ndarray):
        n = self.info().n_points
        if point_idx_offset + n > hssoa.capacity:
            raise RuntimeError("transfer_from_device() HSSOA capacity")
        out = np.zeros((hssoa.nlines, max(n, 1)), np.float32)
        err = C.c_uint32(0)
        dis = C.c_int32(0)
        self._chk(self.lib.ice_download(self.h, abi.ptr(out), out.shape[1], hssoa.nlines, abi.ptr(boundary_forces),
                                        C.byref(err), C.byref(dis)), "transfer_from_device")
        hssoa.host_buffer[:, point_idx_offset: point_idx_offset + n] = out[:, :n]
        self.nPts_partition = n
        self.error_code = err.value
        self.disabled_points_count = dis.value

    def reset_grid(self):
        self._chk(self.lib.ice_reset_grid(self.h), "cuda_reset_grid error")

    def clear_force_accumulator(self):
        self._chk(self.lib.ice_clear_force_accumulator(self.h), "clear_force_accumulator()")

    def p2g(self):
        self._chk(self.lib.ice_p2g(self.h), "p2g kernel")

    def update_nodes(self, simulation_time: float, windSpeed: float = 0.0, windAngle: float = 0.0):
        self._chk(self.lib.ice_update_nodes(self.h, simulation_time, windSpeed, windAngle), "update_nodes")

    def g2p(self, recordPQ: bool, enablePointTransfer: bool = False, applyGlensLaw: int = 0):
        self._chk(self.lib.ice_g2p(self.h, int(recordPQ), int(enablePointTransfer), int(applyGlensLaw)),
                  "g2p kernel")

    def reset_timings(self):
        self._chk(self.lib.ice_reset_timings(self.h), "reset_timings")
        for k in ("timing_10_P2GAndHalo", "timing_20_acceptHalo", "timing_30_updateGrid", "timing_40_G2P",
                  "timing_60_ptsSent", "timing_70_ptsAccepted", "timing_stepTotal"):
            setattr(self, k, 0.0)

    def record_timings(self, enablePointTransfer: bool = False):
        out = (C.c_double * 8)()
        self._chk(self.lib.ice_get_timings(self.h, out), "record_timings")
        self.timing_10_P2GAndHalo = out[1]
        self.timing_30_updateGrid = out[2]
        self.timing_40_G2P = out[3]
        self.timing_stepTotal = out[4]

    def normalize_timings(self, cycles: int):
        coeff = 1000.0 / cycles
        for k in ("timing_10_P2GAndHalo", "timing_20_acceptHalo", "timing_30_updateGrid", "timing_40_G2P",
                  "timing_60_ptsSent", "timing_70_ptsAccepted", "timing_stepTotal"):
            setattr(self, k, getattr(self, k) * coeff)


class GPU_Implementation5:
    """gpu_implementation5.h:23-68 with one partition."""

    def __init__(self, prms: SimParams, nlines: int | None = None):
        self.prms = prms
        GPU_Partition.prms = prms
        self.partitions: list[GPU_Partition] = []
        nl = nlines or (abi.N_LINES_MAX if prms.flags & abi.FLAG_COHESION_LINE else abi.N_LINES)
        self.hssoa = HostSideSOA(nl)
        self.error_code = 0
        N = prms.GridXTotal * prms.GridYTotal
        self.grid_status_buffer = np.full(N, abi.GRID_STATUS_WATER, np.uint8)
        self.grid_boundary_normals = np.zeros(0, np.float32)
        self.grid_boundary_forces = np.zeros(2 * N, np.float32)
        self.current = None
        self.wind_stress = None
        self.transfer_completion_callback = None

    def allocate_host_arrays_grid(self):
        N = self.prms.GridXTotal * self.prms.GridYTotal
        self.grid_status_buffer = np.full(N, abi.GRID_STATUS_WATER, np.uint8)
        self.grid_boundary_normals = np.zeros(2 * N, np.float32)
        self.grid_boundary_forces = np.zeros(2 * N, np.float32)

    def allocate_host_arrays_points(self, nPtsInitial: int):
        self.hssoa.Allocate(nPtsInitial)

    def initialize(self, device: int = 0):
        self.partitions = [GPU_Partition()]
        self.partitions[0].initialize(device, 0)

    def split_hssoa_into_partitions(self):
        p = self.partitions[0]
        p.nPts_partition = self.hssoa.size
        p.GridX_partition = self.prms.GridXTotal
        p.disabled_points_count = 0

    def allocate_device_arrays(self):
        p = self.partitions[0]
        p.allocate(self.hssoa.capacity, p.GridX_partition)

    def transfer_to_device(self):
        p = self.partitions[0]
        p.transfer_points_from_soa_to_device(self.hssoa, 0)
        p.transfer_grid_data_to_device(self)

    def transfer_wind_and_current_data_to_device(self):
        for p in self.partitions:
            p.update_current_field(self.current, self.wind_stress)

    def transfer_from_device(self, elapsed_cycles: int):
        offset = 0
        for p in self.partitions:
            p.transfer_from_device(self.hssoa, offset, self.grid_boundary_forces)
            offset += p.nPts_partition
            if p.error_code:
                self.error_code = p.error_code
        self.hssoa.size = offset
        # boundary-force normalisation (gpu_implementation5.cpp:219-221)
        self.grid_boundary_forces /= (elapsed_cycles * self.prms.InitialTimeStep)
        if self.transfer_completion_callback:
            self.transfer_completion_callback()

    def synchronize(self):
        for p in self.partitions:
            p._chk(p.lib.ice_synchronize(p.h), "synchronize")

    def update_constants(self):
        self.error_code = 0
        for p in self.partitions:
            p.update_constants()

    def reset_grid(self):
        for p in self.partitions:
            p.reset_grid()

    def reset_timings(self):
        for p in self.partitions:
            p.reset_timings()

    def clear_force_accumulator(self):
        for p in self.partitions:
            p.clear_force_accumulator()

    def p2g(self):
        self.partitions[0].p2g()

    def update_nodes(self, simulation_time: float, windSpeed: float = 0.0, windAngle: float = 0.0):
        for p in self.partitions:
            p.update_nodes(simulation_time, windSpeed, windAngle)

    def g2p(self, recordPQ: bool, enablePointTransfer: bool = False, applyGlensLaw: int = 0):
        for p in self.partitions:
            p.g2p(recordPQ, enablePointTransfer, applyGlensLaw)

    def record_timings(self, enablePointTransfer: bool = False):
        for p in self.partitions:
            p.record_timings(enablePointTransfer)


class ForcingMixin:
    def upload_wind_frames(self, wind0, wind1, t0: float, t1: float, d: int = 0):
        """Time-varying wind stress of domain d: frames [2][gx*gy] (Pa) valid at
        t0 < t1, interpolated linearly in time by every later substep
        (ice_upload_wind_frames)."""
        a = np.ascontiguousarray(wind0, np.float32)
        b = np.ascontiguousarray(wind1, np.float32)
        self._chk(self.lib.ice_upload_wind_frames(self.h, d, abi.ptr(a), abi.ptr(b), float(t0), float(t1)),
                  "ice_upload_wind_frames")


class SurrogateMixin(ForcingMixin):
    """The learned plastic projection's entry points (ice_mlp_*) on a handle."""

    def mlp_generate(self, n: int, seed: int = 1, sampling: "abi.IceMlpSampling | None" = None, copy: bool = False):
        x = np.zeros((n, 4), np.float32) if copy else None
        t = np.zeros((n, 3), np.float32) if copy else None
        failed = C.c_uint32(0)
        self._chk(self.lib.ice_mlp_generate(self.h, n, seed, C.byref(sampling) if sampling is not None else None,
                                            abi.ptr(x), abi.ptr(t), C.byref(failed)), "ice_mlp_generate")
        return x, t, failed.value

    def mlp_record(self, capacity: int, keep_fraction: float = 1.0):
        """Record ridge-capped plastic trial states of the following substeps
        (capacity 0 stops and frees)."""
        self._chk(self.lib.ice_mlp_record(self.h, capacity, keep_fraction), "ice_mlp_record")

    def mlp_record_keep(self, keep_fraction: float):
        """Change the kept fraction of the open recording."""
        self._chk(self.lib.ice_mlp_record_keep(self.h, keep_fraction), "ice_mlp_record_keep")

    def mlp_record_count(self):
        """(states seen, rows stored)."""
        seen, stored = C.c_uint64(0), C.c_uint32(0)
        self._chk(self.lib.ice_mlp_record_count(self.h, C.byref(seen), C.byref(stored)), "ice_mlp_record_count")
        return int(seen.value), int(stored.value)

    def mlp_records(self, n_max: int) -> np.ndarray:
        """The stored trial-state rows [n][4] (log J, D, Jp_inv, log(c / 1e5))."""
        x = np.zeros((max(n_max, 1), 4), np.float32)
        n = C.c_uint32(0)
        self._chk(self.lib.ice_mlp_records(self.h, abi.ptr(x), n_max, C.byref(n)), "ice_mlp_records")
        return x[:n.value]

    def train_on_recorded(self, first_step: int, record_steps: int = 16, pairs: int = 10_000_000,
                          train_steps=(8000, 6000, 6000), lrs=(2e-3, 5e-4, 1e-4), batch: int = 8192):
        """Train the learned projection on the handle's own workload: record
        about `pairs` ridge-capped plastic trial states (hash-sampled) over
        `record_steps` substeps, draw `pairs` training pairs from the records
        (targets from the exact FP64 map) and train with Adam.  The first
        substep only counts the states; before each later one the kept
        fraction is set from the last substep's count so that the records
        spread over the whole window even where the count grows as the run
        develops.  Runs 1 + record_steps substeps from first_step.  Returns a
        report dict (states seen, kept fractions, loss, seconds)."""
        import time
        cap, target = int(1.2 * pairs) + 1024, 1.1 * pairs
        self.mlp_record(cap, 0.0)                   # the first substep counts only
        keeps, seen_prev, seen1 = [], 0, None
        for k in range(record_steps + 1):
            self.run(first_step + k, 1, graph=False)
            seen, stored = self.mlp_record_count()
            last, seen_prev = seen - seen_prev, seen
            seen1 = last if seen1 is None else seen1
            rem = record_steps - k
            if rem:
                keeps.append(min(1.0, max(target - stored, 0.0) / max(last * rem, 1)))
                self.mlp_record_keep(keeps[-1])
        seen, stored = self.mlp_record_count()
        t0 = time.perf_counter()
        _, _, failed = self.mlp_generate_recorded(pairs, seed=3)
        self.mlp_record(0, 0.0)
        self.mlp_init(seed=21)
        loss = np.concatenate([self.mlp_train(st, batch=batch, lr=lr, seed=5 + k)
                               for k, (st, lr) in enumerate(zip(train_steps, lrs))])
        return {"states_seen_per_substep": seen1, "states_seen": seen, "states_stored": stored,
                "keep_fractions": [round(k, 4) for k in keeps], "pairs": pairs, "pairs_failed": failed, "train_steps": int(loss.shape[0]),
                "seconds": time.perf_counter() - t0, "loss_first": float(loss[0]),
                "loss_last": float(loss[-100:].mean())}

    def mlp_generate_recorded(self, n: int, seed: int = 1, copy: bool = False):
        x = np.zeros((n, 4), np.float32) if copy else None
        t = np.zeros((n, 3), np.float32) if copy else None
        failed = C.c_uint32(0)
        self._chk(self.lib.ice_mlp_generate_recorded(self.h, n, seed, abi.ptr(x), abi.ptr(t), C.byref(failed)),
                  "ice_mlp_generate_recorded")
        return x, t, failed.value

    def mlp_norm(self) -> "abi.IceMlpNorm":
        nrm = abi.IceMlpNorm()
        self._chk(self.lib.ice_mlp_get_norm(self.h, C.byref(nrm)), "ice_mlp_get_norm")
        return nrm

    def mlp_set_norm(self, nrm: "abi.IceMlpNorm"):
        self._chk(self.lib.ice_mlp_set_norm(self.h, C.byref(nrm)), "ice_mlp_set_norm")

    def mlp_init(self, seed: int = 1):
        self._chk(self.lib.ice_mlp_init_params(self.h, seed), "ice_mlp_init_params")

    def mlp_params(self) -> np.ndarray:
        p = np.zeros(abi.MLP_PARAMS, np.float32)
        self._chk(self.lib.ice_mlp_get_params(self.h, abi.ptr(p)), "ice_mlp_get_params")
        return p

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (abi.MLP_PARAMS,)
        self._chk(self.lib.ice_mlp_set_params(self.h, abi.ptr(p)), "ice_mlp_set_params")

    def mlp_train(self, steps: int, batch: int = 8192, lr: float = 1e-3, seed: int = 7) -> np.ndarray:
        loss = np.zeros(max(steps, 1), np.float32)
        self._chk(self.lib.ice_mlp_train(self.h, steps, batch, lr, seed, abi.ptr(loss)), "ice_mlp_train")
        return loss[:steps]

    def mlp_forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        y = np.zeros((x.shape[0], 3), np.float32)
        self._chk(self.lib.ice_mlp_forward(self.h, abi.ptr(x), x.shape[0], abi.ptr(y)), "ice_mlp_forward")
        return y

    def mlp_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 3)
        g = np.zeros(abi.MLP_PARAMS, np.float32)
        loss = C.c_float(0)
        self._chk(self.lib.ice_mlp_grad(self.h, abi.ptr(x), abi.ptr(t), x.shape[0], abi.ptr(g), C.byref(loss)),
                  "ice_mlp_grad")
        return float(loss.value), g

    def mlp_eval(self, x, tol: float = 1e-4) -> np.ndarray:
        """[n][6]: surrogate p', q'; exact p', q'; yield violation; accepted."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        out = np.zeros((x.shape[0], 6), np.float32)
        self._chk(self.lib.ice_mlp_eval(self.h, abi.ptr(x), x.shape[0], tol, abi.ptr(out)), "ice_mlp_eval")
        return out

    def mlp_deploy(self, enable: bool = True, tol: float = 1e-4):
        self._chk(self.lib.ice_mlp_deploy(self.h, int(enable), tol), "ice_mlp_deploy")

    def mlp_counts(self):
        c = (C.c_uint64 * 2)()
        self._chk(self.lib.ice_mlp_counts(self.h, c), "ice_mlp_counts")
        return int(c[0]), int(c[1])



class Simulation(SurrogateMixin):
    """A scene on one GPU through the C-ABI: upload, run, download."""

    def __init__(self, scene, device: int = 0, *, flags: int | None = None, tile=None, capacity: int | None = None,
                 options: dict | None = None):
        self.lib = abi.load()
        self.scene = scene
        prm = scene.params if flags is None else scene.params.replace(flags=flags)
        self.params = prm
        self.cp = prm.to_c()
        self.h = C.c_void_p()
        abi.check(self.lib.ice_create(C.byref(self.cp), device, C.byref(self.h)), None, "ice_create")
        if tile is not None:
            self._chk(self.lib.ice_set_tile(self.h, tile[0], tile[1]), "ice_set_tile")
        for name, value in (options or {}).items():
            self.set_option(name, value)
        self.nlines = scene.nlines
        self._chk(self.lib.ice_allocate(self.h, capacity or scene.n), "ice_allocate")
        self._chk(self.lib.ice_upload_grid(self.h, abi.ptr(np.ascontiguousarray(scene.status)), None),
                  "ice_upload_grid")
        cur = None if scene.current is None else np.ascontiguousarray(scene.current, np.float32)
        wnd = None if scene.wind is None else np.ascontiguousarray(scene.wind, np.float32)
        self._chk(self.lib.ice_upload_forcing(self.h, abi.ptr(cur), abi.ptr(wnd)), "ice_upload_forcing")
        soa = np.ascontiguousarray(scene.soa, np.float32)
        self._chk(self.lib.ice_upload_points(self.h, abi.ptr(soa), soa.shape[1], soa.shape[1], soa.shape[0]),
                  "ice_upload_points")

    def _chk(self, rc, what):
        abi.check(rc, self.h, what)

    def reset_points(self):
        """Upload the scene's initial points again (the state restarts)."""
        soa = np.ascontiguousarray(self.scene.soa, np.float32)
        self._chk(self.lib.ice_upload_points(self.h, abi.ptr(soa), soa.shape[1], soa.shape[1], soa.shape[0]),
                  "ice_upload_points")

    def close(self):
        if self.h:
            self.lib.ice_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, name: str, value: int):
        self._chk(self.lib.ice_set_option(self.h, name.encode(), int(value)), f"ice_set_option({name})")

    def substep(self, t: float, record_pq: bool = False):
        self._chk(self.lib.ice_substep(self.h, t, int(record_pq)), "ice_substep")

    def run(self, first_step: int, nsteps: int, graph: bool = True):
        self._chk(self.lib.ice_run(self.h, first_step, nsteps, int(graph)), "ice_run")

    def sync(self):
        self._chk(self.lib.ice_synchronize(self.h), "ice_synchronize")

    def info(self) -> abi.IceInfo:
        i = abi.IceInfo()
        self._chk(self.lib.ice_get_info(self.h, C.byref(i)), "ice_get_info")
        return i

    def download(self):
        """(points [nlines, n], forces [2, N], error word, disabled count)."""
        n = self.info().n_points
        out = np.zeros((self.nlines, max(n, 1)), np.float32)
        N = self.params.GridXTotal * self.params.GridYTotal
        forces = np.zeros((2, N), np.float32)
        err, dis = C.c_uint32(0), C.c_int32(0)
        self._chk(self.lib.ice_download(self.h, abi.ptr(out), out.shape[1], self.nlines, abi.ptr(forces),
                                        C.byref(err), C.byref(dis)), "ice_download")
        return out[:, :n], forces, err.value, dis.value

    def grid(self):
        N = self.params.GridXTotal * self.params.GridYTotal
        mpp = np.zeros((3, N), np.float32)
        vel = np.zeros((2, N), np.float32)
        self._chk(self.lib.ice_download_grid(self.h, abi.ptr(mpp), abi.ptr(vel)), "ice_download_grid")
        return mpp, vel

    def upload_point_colors(self, rgb):
        """Point colours 0xRRGGBB by original point index (point_colors_rgb)."""
        rgb = np.ascontiguousarray(rgb, np.uint32).reshape(-1)
        self._chk(self.lib.ice_upload_point_colors(self.h, abi.ptr(rgb), rgb.shape[0]), "ice_upload_point_colors")

    def prepare_frame_arrays(self, elapsed_cycles: int = 0,
                             want=("count", "fields", "rgb", "mass_mask", "forces_per_region")) -> dict:
        """SnapshotManager::PrepareFrameArrays (snapshotmanager.cpp:623-765) on
        the device: the ten fields [N] by name (abi.FRAME_FIELD_NAMES; Jpinv
        is Jp_inv - 1), count, rgb [N][3], mass_mask, forces_per_region
        [100][2] (those named in `want`).  Sorts the points by cell first."""
        N = self.params.GridXTotal * self.params.GridYTotal
        count = np.zeros(N, np.uint8) if "count" in want else None
        fields = np.zeros((abi.FRAME_FIELDS, N), np.float32) if "fields" in want else None
        rgb = np.zeros((N, 3), np.uint8) if "rgb" in want else None
        mask = np.zeros(N, np.uint8) if "mass_mask" in want else None
        fpr = np.zeros((abi.MAX_REGIONS, 2), np.float32) if "forces_per_region" in want else None
        fa = abi.IceFrameArrays(abi.ptr(count), abi.ptr(fields), abi.ptr(rgb), abi.ptr(mask), abi.ptr(fpr))
        self._chk(self.lib.ice_prepare_frame_arrays(self.h, elapsed_cycles, C.byref(fa)), "ice_prepare_frame_arrays")
        out = {name: fields[k] for k, name in enumerate(abi.FRAME_FIELD_NAMES)} if fields is not None else {}
        out.update(count=count, rgb=rgb, mass_mask=mask, forces_per_region=fpr)
        return out

    def sort(self, want_perm: bool = False):
        n0 = self.info().n_points
        perm = np.zeros(max(n0, 1), np.uint32) if want_perm else None
        nout = C.c_uint32(0)
        self._chk(self.lib.ice_sort(self.h, C.byref(nout), abi.ptr(perm)), "ice_sort")
        return perm[: nout.value] if want_perm else nout.value

    def timings(self):
        out = (C.c_double * 8)()
        self._chk(self.lib.ice_get_timings(self.h, out), "ice_get_timings")
        return list(out)

    def enable_timings(self, on: bool = True):
        self._chk(self.lib.ice_enable_timings(self.h, int(on)), "ice_enable_timings")

    def reset_timings(self):
        self._chk(self.lib.ice_reset_timings(self.h), "ice_reset_timings")

    def autotune(self, nsteps: int = 3):
        tw, th, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        self._chk(self.lib.ice_autotune(self.h, nsteps, C.byref(tw), C.byref(th), C.byref(ms)), "ice_autotune")
        return (tw.value, th.value), ms.value




class BatchSimulation(SurrogateMixin):
    """Several independent scenes in one batched handle (ice_create_batch):
    one launch per phase over all domains' tiles, HIP-graph replays."""

    def __init__(self, scene_list, device: int = 0, *, flags: int | None = None, tile=None, capacity_factor=1.0,
                 options: dict | None = None):
        self.lib = abi.load()
        self.scenes = list(scene_list)
        p0 = self.scenes[0].params
        prm = p0 if flags is None else p0.replace(flags=flags)
        for s in self.scenes:
            if s.params.ParticleVolume != p0.ParticleVolume or s.nlines != self.scenes[0].nlines:
                raise ValueError("batched domains share material constants, ParticleVolume and the line count")
        self.params = prm
        self.cp = prm.to_c()
        nd = len(self.scenes)
        gx = np.array([s.params.GridXTotal for s in self.scenes], np.int32)
        gy = np.array([s.params.GridYTotal for s in self.scenes], np.int32)
        cap = np.array([max(1, int(s.n * capacity_factor)) for s in self.scenes], np.uint32)
        self.h = C.c_void_p()
        abi.check(self.lib.ice_create_batch(C.byref(self.cp), device, nd, abi.ptr(gx), abi.ptr(gy), abi.ptr(cap),
                                            C.byref(self.h)), None, "ice_create_batch")
        if tile is not None:
            self._chk(self.lib.ice_set_tile(self.h, tile[0], tile[1]), "ice_set_tile")
        for name, value in (options or {}).items():
            self._chk(self.lib.ice_set_option(self.h, name.encode(), int(value)), f"ice_set_option({name})")
        self.nlines = self.scenes[0].nlines
        self.reset_points()

    def reset_points(self):
        """Upload every domain's initial points, status and forcing (the state restarts)."""
        for d, s in enumerate(self.scenes):
            soa = np.ascontiguousarray(s.soa, np.float32)
            st = np.ascontiguousarray(s.status, np.uint8)
            cur = None if s.current is None else np.ascontiguousarray(s.current, np.float32)
            wnd = None if s.wind is None else np.ascontiguousarray(s.wind, np.float32)
            self._chk(self.lib.ice_upload_domain(self.h, d, abi.ptr(soa), soa.shape[1], soa.shape[1], soa.shape[0],
                                                 abi.ptr(st), abi.ptr(cur), abi.ptr(wnd)), f"ice_upload_domain {d}")

    def _chk(self, rc, what):
        abi.check(rc, self.h, what)

    def close(self):
        if self.h:
            self.lib.ice_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, first_step: int, nsteps: int, graph: bool = True):
        self._chk(self.lib.ice_run(self.h, first_step, nsteps, int(graph)), "ice_run")

    def substep(self, t: float, record_pq: bool = False):
        self._chk(self.lib.ice_substep(self.h, t, int(record_pq)), "ice_substep")

    def sync(self):
        self._chk(self.lib.ice_synchronize(self.h), "ice_synchronize")

    def info(self, d: int = 0) -> abi.IceInfo:
        i = abi.IceInfo()
        self._chk(self.lib.ice_get_domain_info(self.h, d, C.byref(i)), "ice_get_domain_info")
        return i

    def n_points(self) -> int:
        return sum(self.info(d).n_points for d in range(len(self.scenes)))

    def download(self, d: int):
        n = self.info(d).n_points
        out = np.zeros((self.nlines, max(n, 1)), np.float32)
        N = self.scenes[d].params.GridXTotal * self.scenes[d].params.GridYTotal
        forces = np.zeros((2, N), np.float32)
        err, dis = C.c_uint32(0), C.c_int32(0)
        self._chk(self.lib.ice_download_domain(self.h, d, abi.ptr(out), out.shape[1], self.nlines, abi.ptr(forces),
                                               C.byref(err), C.byref(dis)), "ice_download_domain")
        return out[:, :n], forces, err.value, dis.value

    def autotune(self, nsteps: int = 3):
        tw, th, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        self._chk(self.lib.ice_autotune(self.h, nsteps, C.byref(tw), C.byref(th), C.byref(ms)), "ice_autotune")
        return (tw.value, th.value), ms.value

    def enable_timings(self, on: bool = True):
        self._chk(self.lib.ice_enable_timings(self.h, int(on)), "ice_enable_timings")

    def reset_timings(self):
        self._chk(self.lib.ice_reset_timings(self.h), "ice_reset_timings")

    def timings(self):
        out = (C.c_double * 8)()
        self._chk(self.lib.ice_get_timings(self.h, out), "ice_get_timings")
        return list(out)
