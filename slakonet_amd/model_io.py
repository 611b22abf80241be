"""Model save / load in the reference's on-disk layouts.

Formats (slakonet/optim.py:204-625):
* "state_dict": directory <path without suffix>/ holding model_state.pt
  (the module state dict: skf_optimizers.{A-B}.h_params.{l1-l2} ...),
  metadata.json and skf_data.pt ({pair: {skf_dict, original_h_params,
  original_s_params}}) (:226-266, load :309-403);
* "compact": one <path>.pt with {metadata (ultra_compact: True),
  trained_parameters (state dict), skf_metadata ({pair: skf_dict without
  hamiltonian / overlap})} (:485-534, load :536-625); on load the trained
  values are also the constraint originals, as the file keeps no others;
* "universal_params": <path>_universal_params.pt with {pair: {h_params,
  s_params, skf_dict}} (:277-299, load :405-421).
"full_model" (a pickled module) is not offered: files are read with
torch.load(weights_only=True) only, which executes nothing from the file.
"""
from __future__ import annotations

import json
from pathlib import Path

import torch

from .units import SYMBOL_TO_Z


def _pairs_to_z(key: str):
    a, b = key.split("-")
    return SYMBOL_TO_Z[a], SYMBOL_TO_Z[b]


def _metadata(model, compact=False):
    md = {
        "skf_directory": None if model.skf_directory is None else str(model.skf_directory),
        "elements_in_system": sorted(model.elements_in_system),
        "element_pairs": [list(p) for p in sorted(model.element_pairs)],
        "available_pairs": list(model.skf_optimizers.keys()),
        "class_name": "MultiElementSkfParameterOptimizer",
    }
    if compact:
        md["ultra_compact"] = True
    return md


def _cpu_state(model):
    return {k: v.detach().cpu() for k, v in model.state_dict().items()}


def save_model(model, save_path, method="state_dict"):
    save_path = Path(save_path)
    save_path.parent.mkdir(parents=True, exist_ok=True)
    if method == "state_dict":
        d = save_path.with_suffix("")
        d.mkdir(parents=True, exist_ok=True)
        torch.save(_cpu_state(model), d / "model_state.pt")
        with open(d / "metadata.json", "w") as f:
            json.dump(_metadata(model), f, indent=2)
        skf_data = {}
        for key, opt in model.skf_optimizers.items():
            skf_data[key] = {
                "skf_dict": opt.skf_dict,
                "original_h_params": {k: v.tolist() for k, v in opt.original_h_params.items()},
                "original_s_params": {k: v.tolist() for k, v in opt.original_s_params.items()},
            }
        torch.save(skf_data, d / "skf_data.pt")
        return d
    if method == "compact":
        f = save_path.with_suffix(".pt")
        meta = {}
        for key, opt in model.skf_optimizers.items():
            sd = dict(opt.skf_dict)
            sd.pop("hamiltonian", None)
            sd.pop("overlap", None)
            meta[key] = sd
        torch.save({"metadata": _metadata(model, compact=True), "trained_parameters": _cpu_state(model),
                    "skf_metadata": meta}, f)
        return f
    if method == "universal_params":
        f = save_path.parent / f"{save_path.name}_universal_params.pt"
        out = {}
        for key, opt in model.skf_optimizers.items():
            out[key] = {"h_params": {k: v.detach().cpu().clone() for k, v in opt.h_params.items()},
                        "s_params": {k: v.detach().cpu().clone() for k, v in opt.s_params.items()},
                        "skf_dict": opt.skf_dict}
        torch.save(out, f)
        return f
    if method == "full_model":
        raise NotImplementedError("full_model pickles the module; use state_dict, compact or universal_params")
    raise ValueError(f"Unknown save method: {method}")


def _split_state(state, key):
    h, s = {}, {}
    for name, v in state.items():
        if name.startswith(f"skf_optimizers.{key}.h_params."):
            h[name.rsplit(".", 1)[1]] = v
        elif name.startswith(f"skf_optimizers.{key}.s_params."):
            s[name.rsplit(".", 1)[1]] = v
    return h, s


def load_model(cls, load_path, method="state_dict", skf_directory=None):
    load_path = Path(load_path)
    if method == "state_dict":
        d = load_path.with_suffix("") if load_path.is_file() else load_path
        with open(d / "metadata.json") as f:
            md = json.load(f)
        skf_data = torch.load(d / "skf_data.pt", map_location="cpu", weights_only=True)
        dicts = {}
        for key, entry in skf_data.items():
            sd = dict(entry["skf_dict"])
            sd["hamiltonian"] = entry["original_h_params"]
            sd["overlap"] = entry["original_s_params"]
            dicts[_pairs_to_z(key)] = sd
        model = cls(skf_dicts=dicts)
        model.skf_directory = md.get("skf_directory")
        model.load_state_dict(torch.load(d / "model_state.pt", map_location="cpu", weights_only=True))
        return model
    if method == "compact":
        data = torch.load(load_path.with_suffix(".pt"), map_location="cpu", weights_only=True)
        if not data["metadata"].get("ultra_compact", False):
            raise ValueError("This is not an ultra-compact model file")
        state = data["trained_parameters"]
        dicts = {}
        for key in data["metadata"]["available_pairs"]:
            sd = dict(data["skf_metadata"][key])
            sd["hamiltonian"], sd["overlap"] = _split_state(state, key)
            dicts[_pairs_to_z(key)] = sd
        model = cls(skf_dicts=dicts)
        model.skf_directory = data["metadata"].get("skf_directory")
        model.load_state_dict(state)
        return model
    if method == "universal_params":
        f = load_path.parent / f"{load_path.name}_universal_params.pt"
        from .optim import _load_universal
        return cls(skf_dicts=_load_universal(str(f)))
    if method == "full_model":
        raise NotImplementedError("full_model files are pickles and are not loaded")
    raise ValueError(f"Unknown load method: {method}")
