"""Learned pose proposal (include/rln_proposal.h): ctypes wrapper and the seeding of coarse orientation lists.

The network, its training and its inference run in librln_estep.so; this module only moves arguments. The
product path has no CPU fallback: the calls raise when the library or the device is missing.
"""
import ctypes as C

import numpy as np

from . import (ProposalConfig, ProposalGen, _check, _float_pairs, _ptr, _vp, lib)


class PoseProposal:
    """Direction x psi bin classifier seeding the E-step's coarse search (rln_proposal)."""

    def __init__(self, window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=1, device=0):
        self.cfg = ProposalConfig(window, channels, healpix_order, n_psi_bins, seed)
        self._pending = None
        self._h = _vp()
        _check(lib.rln_proposal_create(device, C.byref(self.cfg), C.byref(self._h)))
        n = C.c_longlong()
        _check(lib.rln_proposal_num_params(self._h, C.byref(n)))
        self.n_params = n.value
        nb = C.c_int()
        _check(lib.rln_proposal_num_bins(self._h, C.byref(nb)))
        self.n_bins = nb.value

    # parameters -------------------------------------------------------------------------------
    def get_params(self):
        out = np.zeros(self.n_params, np.float32)
        _check(lib.rln_proposal_get_params(self._h, _ptr(out, C.c_float)))
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.n_params
        _check(lib.rln_proposal_set_params(self._h, _ptr(w, C.c_float)))

    # training set -----------------------------------------------------------------------------
    def clear_training(self):
        _check(lib.rln_proposal_clear_training(self._h))

    def generate(self, projector, n, ori_size, pixel_size, sigma2_noise, defocus=(8000.0, 25000.0),
                 max_astigmatism=0.0, offset_max=3.0, voltage=300.0, cs=2.7, q0=0.1, first=0):
        """Adds n projections of the reference at random poses (with CTF and noise); returns their poses."""
        s2 = np.ascontiguousarray(sigma2_noise, np.float64)
        g = ProposalGen(ori_size, pixel_size, voltage, cs, q0, defocus[0], defocus[1], max_astigmatism,
                        offset_max, _ptr(s2, C.c_double))
        poses = np.zeros((n, 3))
        _check(lib.rln_proposal_generate(self._h, projector._h, C.byref(g), n, first, _ptr(poses, C.c_double)))
        return poses

    def add_images(self, fimg, fctf, current_size, poses, on_device=False):
        """Adds particle images (and CTFs, or None) labelled with poses (rot, tilt, psi) -- e.g. max-likelihood
        assignments."""
        poses = np.ascontiguousarray(poses, np.float64)
        n = len(poses)
        a, keep = self._images_arg(fimg, on_device)
        c, keep_c = self._images_arg(fctf, on_device)
        _check(lib.rln_proposal_add_images(self._h, n, a, c, current_size, 1 if on_device else 0,
                                           _ptr(poses, C.c_double)))
        del keep, keep_c

    def training_size(self):
        n = C.c_longlong()
        _check(lib.rln_proposal_training_size(self._h, C.byref(n)))
        return n.value

    def train(self, epochs=1, batch=256, lr=1e-3):
        loss = C.c_double()
        _check(lib.rln_proposal_train(self._h, epochs, batch, lr, C.byref(loss)))
        return loss.value

    # inference --------------------------------------------------------------------------------
    @staticmethod
    def _images_arg(fimg, on_device):
        if fimg is None:
            return None, None
        if on_device:
            return _vp(int(fimg)), None
        fa = _float_pairs(fimg)
        return _vp(fa.ctypes.data), fa

    def predict(self, fimg, fctf, n, current_size, topk=16, on_device=False):
        """Top-k bins and probabilities of n particle images (fctf: their CTFs for phase flipping, or None)."""
        bins = np.zeros((n, topk), np.int32)
        probs = np.zeros((n, topk), np.float32)
        a, keep = self._images_arg(fimg, on_device)
        c, keep_c = self._images_arg(fctf, on_device)
        _check(lib.rln_proposal_predict(self._h, n, a, c, current_size, 1 if on_device else 0, topk,
                                        _ptr(bins, C.c_int), _ptr(probs, C.c_float)))
        del keep, keep_c
        return bins, probs

    def predict_launch(self, fimg, fctf, n, current_size, topk=16, on_device=False):
        """Enqueues a prediction on the proposal's stream (rln_proposal_predict_launch); predict_finish returns it."""
        a, keep = self._images_arg(fimg, on_device)
        c, keep_c = self._images_arg(fctf, on_device)
        _check(lib.rln_proposal_predict_launch(self._h, n, a, c, current_size, 1 if on_device else 0, topk))
        self._pending = (n, topk, keep, keep_c)  # host inputs stay alive until the finish

    def predict_finish(self):
        if self._pending is None:
            _check(lib.rln_proposal_predict_finish(self._h, None, None))  # raises: no prediction in flight
        n, topk, _, _ = self._pending
        bins = np.zeros((n, topk), np.int32)
        probs = np.zeros((n, topk), np.float32)
        try:
            _check(lib.rln_proposal_predict_finish(self._h, _ptr(bins, C.c_int), _ptr(probs, C.c_float)))
        finally:
            self._pending = None
        return bins, probs

    def pose_bin(self, rot, tilt, psi):
        b = C.c_int()
        _check(lib.rln_proposal_pose_bin(self._h, rot, tilt, psi, C.byref(b)))
        return b.value

    def bin_orientations(self, grid_order, n_psi_grid, b):
        cnt = C.c_int()
        _check(lib.rln_proposal_bin_orientations(self._h, grid_order, n_psi_grid, b, None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, np.int32)
        _check(lib.rln_proposal_bin_orientations(self._h, grid_order, n_psi_grid, b, _ptr(out, C.c_int), cnt.value,
                                                 C.byref(cnt)))
        return out

    def orientation_lists(self, bins, probs, grid_order, n_psi_grid, min_top1=0.2):
        """Coarse orientation lists from the top bins as (offsets[n+1], orientations); a particle whose top-1
        probability is below min_top1 gets an empty list (exhaustive search)."""
        bins = np.ascontiguousarray(bins, np.int32)
        probs = np.ascontiguousarray(probs, np.float32)
        n, topk = bins.shape
        cap = n * topk * len(self.bin_orientations(grid_order, n_psi_grid, 0))
        off = np.zeros(n + 1, np.int32)
        og = np.empty(max(cap, 1), np.int32)  # the library writes the first off[n] entries, the rest is not returned
        _check(lib.rln_proposal_orientation_sets(self._h, n, _ptr(bins, C.c_int), _ptr(probs, C.c_float), topk,
                                                 grid_order, n_psi_grid, min_top1, _ptr(off, C.c_int),
                                                 _ptr(og, C.c_int), cap))
        return off, og[:off[-1]]

    def orientation_sets(self, bins, probs, grid_order, n_psi_grid, min_top1=0.2):
        """orientation_lists as one array per particle."""
        off, og = self.orientation_lists(bins, probs, grid_order, n_psi_grid, min_top1)
        return [og[off[i]:off[i + 1]] for i in range(len(off) - 1)]

    # check hooks ------------------------------------------------------------------------------
    def images(self, fimg, fctf, n, current_size, on_device=False):
        W = self.cfg.window
        out = np.zeros((n, W, W), np.float32)
        a, keep = self._images_arg(fimg, on_device)
        c, keep_c = self._images_arg(fctf, on_device)
        _check(lib.rln_proposal_images(self._h, n, a, c, current_size, 1 if on_device else 0, _ptr(out, C.c_float)))
        del keep, keep_c
        return out

    def forward_backward(self, images, labels):
        images = np.ascontiguousarray(images, np.float32)
        labels = np.ascontiguousarray(labels, np.int32)
        n = len(labels)
        loss = C.c_double()
        grads = np.zeros(self.n_params, np.float32)
        logits = np.zeros((n, self.n_bins), np.float32)
        _check(lib.rln_proposal_forward_backward(self._h, n, _ptr(images, C.c_float), _ptr(labels, C.c_int),
                                                 C.byref(loss), _ptr(grads, C.c_float), _ptr(logits, C.c_float)))
        return loss.value, grads, logits

    def close(self):
        if self._h:
            lib.rln_proposal_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
