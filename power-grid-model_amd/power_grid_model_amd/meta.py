"""Dataset metadata built from the library's PGM_meta_* tables (same construction as the reference's
src/power_grid_model/_core/power_grid_meta.py): numpy structured dtypes with the exported offsets,
null values (NaN for float, minimum integer for int8/int32) and initialize_array."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._lib import handle, lib

_CTYPE = {0: np.dtype("i4"), 1: np.dtype("i1"), 2: np.dtype("f8"), 3: np.dtype(("f8", (3,)))}
_NULL = {0: np.iinfo("i4").min, 1: np.iinfo("i1").min, 2: np.nan, 3: np.nan}


@dataclass
class ComponentMetaData:
    dtype: np.dtype
    dtype_dict: dict
    nans: dict
    nan_scalar: np.ndarray


def _build():
    h = handle().ptr
    out: dict[str, dict[str, ComponentMetaData]] = {}
    for di in range(lib.PGM_meta_n_datasets(h)):
        ds = lib.PGM_meta_get_dataset_by_idx(h, di)
        ds_name = lib.PGM_meta_dataset_name(h, ds).decode()
        comps = {}
        for ci in range(lib.PGM_meta_n_components(h, ds)):
            c = lib.PGM_meta_get_component_by_idx(h, ds, ci)
            name = lib.PGM_meta_component_name(h, c).decode()
            size = lib.PGM_meta_component_size(h, c)
            names, formats, offsets, nans = [], [], [], {}
            for ai in range(lib.PGM_meta_n_attributes(h, c)):
                a = lib.PGM_meta_get_attribute_by_idx(h, c, ai)
                an = lib.PGM_meta_attribute_name(h, a).decode()
                ct = lib.PGM_meta_attribute_ctype(h, a)
                names.append(an)
                formats.append(_CTYPE[ct])
                offsets.append(lib.PGM_meta_attribute_offset(h, a))
                nans[an] = _NULL[ct]
            dd = {"names": names, "formats": formats, "offsets": offsets, "itemsize": size, "aligned": True}
            dtype = np.dtype(dd)
            scalar = np.zeros(1, dtype=dtype)
            for an, v in nans.items():
                scalar[an] = v
            comps[name] = ComponentMetaData(dtype=dtype, dtype_dict=dd, nans=nans, nan_scalar=scalar)
        out[ds_name] = comps
    return out


power_grid_meta_data = _build()


def initialize_array(data_type: str, component_type: str, shape: int | tuple, empty: bool = False) -> np.ndarray:
    """Array of records of a component, filled with null values unless empty=True."""
    if not isinstance(shape, tuple):
        shape = (shape,)
    meta = power_grid_meta_data[data_type][component_type]
    if empty:
        return np.empty(shape, dtype=meta.dtype)
    return np.full(shape, meta.nan_scalar[0], dtype=meta.dtype)
