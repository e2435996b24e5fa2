"""Parameter nodes with the reference's vector API (param/base.py, param/parameter.py,
param/param_context.py).

A model owns a group node whose leaves are its parameters in
``_parameter_order``; a group model links the sub-models' nodes in insertion
order.  The flat vector seen by LM is the depth-first list of unlocked leaves
(param/base.py:181-201).  Values live on the host in float64: the gfx950
render kernels receive them by value in their launch descriptors.

Representation maps (utils/conversions/optimization.py:6-54):
  two-sided (lo, hi): r = tan((v-lo) pi/(hi-lo) - pi/2), v = (atan r + pi/2)(hi-lo)/pi + lo
  lower-only lo     : r = v - 1/(v-lo),  v = (r + lo + sqrt((r-lo)^2 + 4)) / 2
  upper-only hi     : r = v - 1/(v-hi),  v = (r + hi - sqrt((r-hi)^2 + 4)) / 2
  cyclic (lo, hi)   : lo + (x - lo) mod (hi - lo) both ways
"""

from __future__ import annotations

import math
from collections import OrderedDict
from contextlib import contextmanager
from typing import Optional

import numpy as np
import torch

from . import AP_config
from .errors import InvalidParameter

__all__ = ["Parameter_Node", "Param_Unlock", "Param_SoftLimits", "Param_Mask"]


def _t(v):
    return torch.as_tensor(v, dtype=AP_config.ap_dtype, device="cpu")


def boundaries(v, lo, hi):
    v = _t(v)
    if lo is None:
        return v - 1.0 / (v - hi)
    if hi is None:
        return v - 1.0 / (v - lo)
    return torch.tan((v - lo) * math.pi / (hi - lo) - math.pi / 2)


def inv_boundaries(r, lo, hi):
    r = _t(r)
    if lo is None:
        return (r + hi - torch.sqrt((r - hi) ** 2 + 4)) * 0.5
    if hi is None:
        return (r + lo + torch.sqrt((r - lo) ** 2 + 4)) * 0.5
    return (torch.arctan(r) + math.pi / 2) * (hi - lo) / math.pi + lo


def d_inv_boundaries(r, lo, hi):
    """d value / d representation (optimization.py:40-47)."""
    r = _t(r)
    if lo is None:
        return 0.5 - 0.5 * (r - hi) / torch.sqrt((r - hi) ** 2 + 4)
    if hi is None:
        return 0.5 + 0.5 * (r - lo) / torch.sqrt((r - lo) ** 2 + 4)
    return (hi - lo) / (math.pi * (r**2 + 1))


def cyclic_boundaries(v, lo, hi):
    return lo + torch.remainder(_t(v) - lo, hi - lo)


class Parameter_Node:
    """A leaf parameter or a group of nodes (param/parameter.py:18-742)."""

    global_unlock = False

    def __init__(self, name, value=None, limits=(None, None), cyclic=False, locked=False,
                 units="none", uncertainty=None, shape=None, state=None, **kwargs):
        self.name = name
        self._identity = None
        self.nodes: "OrderedDict[str, Parameter_Node]" = OrderedDict()
        self.units = units
        self.cyclic = bool(cyclic)
        self.limits = (None if limits is None or limits[0] is None else float(limits[0]),
                       None if limits is None or limits[1] is None else float(limits[1]))
        self.locked = bool(locked)
        self._value = None
        self._uncertainty = None
        self._mask = None
        self.shape = shape
        if state is not None:
            self.set_state(state)
            return
        if value is not None:
            self.value = value
        if uncertainty is not None:
            self.uncertainty = uncertainty

    # -- structure --------------------------------------------------------
    @property
    def leaf(self):
        return len(self.nodes) == 0

    @property
    def identity(self):
        """param/base.py:124-132: the identity a saved state carried, else id()."""
        return id(self) if self._identity is None else self._identity

    def link(self, node):
        self.nodes[node.name] = node

    def __getitem__(self, key):
        if key in self.nodes:
            return self.nodes[key]
        if isinstance(key, str) and ":" in key:
            head, rest = key.split(":", 1)
            return self.nodes[head][rest]
        raise KeyError(f"Unrecognized key for '{self.name}': {key}")

    def __contains__(self, key):
        return key in self.nodes

    def __iter__(self):
        return filter(lambda n: not n.locked, self.nodes.values())

    def flat(self, include_locked=True):
        """param/base.py:181-201: depth-first leaves (deduplicated by identity)."""
        out = OrderedDict()
        if self.leaf and self._value is not None:
            if (not self.locked) or include_locked or Parameter_Node.global_unlock:
                out[self.identity] = self
            return out
        for node in self.nodes.values():
            if node.locked and not (include_locked or Parameter_Node.global_unlock):
                continue
            if node.leaf:
                if node._value is not None:
                    out[node.identity] = node
            else:
                out.update(node.flat(include_locked))
        return out

    def _leaves(self):
        return list(self.flat(include_locked=False).values())

    # -- leaf value -------------------------------------------------------
    @property
    def value(self):
        if self.leaf:
            return self._value
        return {k: n.value for k, n in self.nodes.items()}

    @value.setter
    def value(self, val):
        if self.locked and not Parameter_Node.global_unlock and self._value is not None:
            raise InvalidParameter(f"{self.name} is locked")
        self._set_val_self(val)

    def _set_val_self(self, val):
        """param/parameter.py:422-455."""
        val = _t(val).clone()
        if self.shape is not None:
            val = val.reshape(self.shape)
        else:
            self.shape = tuple(val.shape)
        if self.cyclic:
            self._value = cyclic_boundaries(val, *self.limits)
            return
        lo, hi = self.limits
        if lo is not None and not torch.all(val > lo):
            raise InvalidParameter(f"{self.name} has lower limit {lo}")
        if hi is not None and not torch.all(val < hi):
            raise InvalidParameter(f"{self.name} has upper limit {hi}")
        self._value = val

    def _soft_set_val_self(self, val):
        """param/parameter.py:457-477: push into the limits instead of raising."""
        val = _t(val).clone()
        if self.shape is not None:
            val = val.reshape(self.shape)
        if self.cyclic:
            self._value = cyclic_boundaries(val, *self.limits)
            return
        lo, hi = self.limits
        if lo is not None:
            val = torch.maximum(val, torch.full_like(val, lo + 1e-3))
        if hi is not None:
            val = torch.minimum(val, torch.full_like(val, hi - 1e-3))
        self._value = val

    @property
    def uncertainty(self):
        return self._uncertainty

    @uncertainty.setter
    def uncertainty(self, u):
        self._uncertainty = None if u is None else _t(u).clone()

    @property
    def mask(self):
        if self._mask is None:
            return torch.ones(self.size, dtype=torch.bool)
        return self._mask

    @property
    def size(self):
        if self.leaf:
            return 0 if self._value is None else int(self._value.numel())
        return sum(n.size for n in self._leaves())

    # -- vector API (param/parameter.py:169-420) ---------------------------
    def vector_values(self):
        if self.leaf:
            return self._value.reshape(-1)[self.mask]
        vec = [n.vector_values() for n in self._leaves()]
        return torch.cat(vec) if vec else torch.zeros(0, dtype=AP_config.ap_dtype)

    def vector_mask(self):
        if self.leaf:
            return self.mask.reshape(-1)
        vec = [n.vector_mask() for n in self._leaves()]
        return torch.cat(vec) if vec else torch.zeros(0, dtype=torch.bool)

    def vector_names(self):
        if self.leaf:
            names = [self.name] if self.size == 1 else [f"{self.name}:{i}" for i in range(self.size)]
            return np.array(names)[self.mask.numpy()]
        vec = [n.vector_names() for n in self._leaves()]
        return np.concatenate(vec) if vec else np.array(())

    def vector_identities(self):
        if self.leaf:
            ids = [str(self.identity)] if self.size == 1 else [f"{self.identity}:{i}" for i in range(self.size)]
            return np.array(ids)[self.mask.numpy()]
        vec = [n.vector_identities() for n in self._leaves()]
        return np.concatenate(vec) if vec else np.array(())

    def vector_uncertainty(self):
        if self.leaf:
            if self._uncertainty is None:
                self._uncertainty = torch.ones_like(self._value)
            return self._uncertainty.reshape(-1)[self.mask]
        vec = [n.vector_uncertainty() for n in self._leaves()]
        return torch.cat(vec) if vec else torch.zeros(0, dtype=AP_config.ap_dtype)

    def _split(self, vec):
        """Yield (leaf, slice of vec) pairs following the masked flat order."""
        vec = _t(vec).reshape(-1)
        k = 0
        for n in self._leaves():
            m = int(n.vector_mask().sum())
            yield n, vec[k:k + m]
            k += m
        if k != len(vec):
            raise InvalidParameter(f"vector of length {len(vec)} does not match {k} parameters")

    def vector_set_values(self, values):
        if self.leaf:
            v = self._value.reshape(-1).clone()
            v[self.mask] = _t(values).reshape(-1)
            self._value = v.reshape(self._value.shape)
            return
        for n, v in self._split(values):
            n.vector_set_values(v)

    def vector_set_uncertainty(self, unc):
        if self.leaf:
            if self._uncertainty is None:
                self._uncertainty = torch.ones_like(self._value)
            u = self._uncertainty.reshape(-1).clone()
            u[self.mask] = _t(unc).reshape(-1)
            self._uncertainty = u.reshape(self._value.shape)
            return
        for n, v in self._split(unc):
            n.vector_set_uncertainty(v)

    def vector_set_mask(self, mask):
        mask = torch.as_tensor(mask, dtype=torch.bool).reshape(-1)
        if self.leaf:
            self._mask = mask
            return
        k = 0
        for n in self._leaves():
            n.vector_set_mask(mask[k:k + n.size])
            k += n.size

    def _leaf_rep_to_val(self, rep):
        if self.cyclic:
            return cyclic_boundaries(rep, *self.limits)
        if self.limits[0] is None and self.limits[1] is None:
            return _t(rep)
        return inv_boundaries(rep, *self.limits)

    def _leaf_val_to_rep(self, val):
        if self.cyclic:
            return cyclic_boundaries(val, *self.limits)
        if self.limits[0] is None and self.limits[1] is None:
            return _t(val)
        return boundaries(val, *self.limits)

    def _leaf_dval_drep(self, rep):
        if self.cyclic or (self.limits[0] is None and self.limits[1] is None):
            return torch.ones_like(_t(rep))
        return d_inv_boundaries(rep, *self.limits)

    def vector_transform_rep_to_val(self, rep):
        if self.leaf:
            return self._leaf_rep_to_val(rep)
        out = [n._leaf_rep_to_val(r) for n, r in self._split(rep)]
        return torch.cat(out) if out else torch.zeros(0, dtype=AP_config.ap_dtype)

    def vector_transform_val_to_rep(self, val):
        if self.leaf:
            return self._leaf_val_to_rep(val)
        out = [n._leaf_val_to_rep(v) for n, v in self._split(val)]
        return torch.cat(out) if out else torch.zeros(0, dtype=AP_config.ap_dtype)

    def vector_dval_drep(self, rep):
        """d value / d representation for every element of the flat vector."""
        if self.leaf:
            return self._leaf_dval_drep(rep)
        out = [n._leaf_dval_drep(r) for n, r in self._split(rep)]
        return torch.cat(out) if out else torch.zeros(0, dtype=AP_config.ap_dtype)

    def vector_representation(self):
        return self.vector_transform_val_to_rep(self.vector_values())

    def vector_set_representation(self, rep):
        self.vector_set_values(self.vector_transform_rep_to_val(rep))

    # -- state --------------------------------------------------------------
    def get_state(self):
        """The keys of param/base.py:134-148 and param/parameter.py:598-630:
        name, identity, locked (when locked), nodes (a group's children), and
        for a leaf value, shape, units, uncertainty, limits (None for an open
        side) and cyclic (when cyclic)."""
        state = {"name": self.name, "identity": self.identity}
        if self.locked:
            state["locked"] = True
        if not self.leaf:
            state["nodes"] = [n.get_state() for n in self.nodes.values()]
        if self._value is not None:
            state["value"] = self._value.detach().cpu().tolist()
        if self.shape is not None and self._value is not None:
            state["shape"] = list(self.shape)
        if self.units is not None:
            state["units"] = self.units
        if self._uncertainty is not None:
            state["uncertainty"] = self._uncertainty.detach().cpu().tolist()
        if self.limits != (None, None):
            state["limits"] = list(self.limits)
        if self.cyclic:
            state["cyclic"] = True
        return state

    def set_state(self, state):
        """param/base.py:150-163 and param/parameter.py:632-645: rebuild this
        node, and its children, from a ``get_state`` dictionary."""
        self.name = state["name"]
        self._identity = state.get("identity")
        self.nodes = OrderedDict()
        for sub in state.get("nodes", []):
            self.link(Parameter_Node(sub["name"], state=sub))
        self.units = state.get("units", "none")
        lim = state.get("limits") or (None, None)
        self.limits = (None if lim[0] is None else float(lim[0]), None if lim[1] is None else float(lim[1]))
        self.cyclic = bool(state.get("cyclic", False))
        self.locked = False
        self._value = self._uncertainty = self._mask = None
        self.shape = tuple(state["shape"]) if state.get("shape") is not None else None
        if state.get("value") is not None:
            if isinstance(state["value"], str):
                raise InvalidParameter(f"{self.name}: linked or functional values ({state['value']}) are not supported")
            self._set_val_self(state["value"])
        if state.get("uncertainty") is not None:
            self.uncertainty = state["uncertainty"]
        self.locked = bool(state.get("locked", False))

    def __repr__(self):
        if self.leaf:
            return f"{self.name}: {None if self._value is None else self._value.tolist()}"
        return f"{self.name}: {{" + ", ".join(repr(n) for n in self.nodes.values()) + "}"


@contextmanager
def Param_Unlock(param: Optional[Parameter_Node] = None):
    """param/param_context.py: temporarily unlock a node (all nodes if None)."""
    if param is None:
        Parameter_Node.global_unlock = True
        try:
            yield
        finally:
            Parameter_Node.global_unlock = False
        return
    old = param.locked
    param.locked = False
    try:
        yield
    finally:
        param.locked = old


@contextmanager
def Param_SoftLimits(param: Parameter_Node):
    """param/param_context.py: setting values pushes them into the limits."""
    old = param._set_val_self
    param._set_val_self = param._soft_set_val_self
    try:
        yield
    finally:
        param._set_val_self = old


@contextmanager
def Param_Mask(param: Parameter_Node, mask=None):
    """param/param_context.py: restrict the flat vector to a mask."""
    old = param.vector_mask().clone()
    if mask is not None:
        param.vector_set_mask(mask)
    try:
        yield
    finally:
        param.vector_set_mask(old)
