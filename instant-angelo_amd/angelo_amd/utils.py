"""Helpers mirroring models/utils.py (chunk_batch, get_activation,
scale_anything, cleanup) and the config access the reference gets from
OmegaConf (attribute access plus .get), without OmegaConf.
"""
from collections import defaultdict

import torch
import torch.nn.functional as F


class Config(dict):
    """dict with attribute access and nested wrapping (OmegaConf-like)."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        for k, v in list(self.items()):
            self[k] = _wrap(v)

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def __setattr__(self, k, v):
        self[k] = _wrap(v)

    def copy(self):
        return Config(super().copy())


def _wrap(v):
    if isinstance(v, Config):
        return v
    if isinstance(v, dict):
        return Config(v)
    if isinstance(v, (list, tuple)):
        return type(v)(_wrap(x) for x in v)
    return v


def to_config(d):
    return _wrap(d)


def config_to_primitive(c):
    if isinstance(c, dict):
        return {k: config_to_primitive(v) for k, v in c.items()}
    if isinstance(c, (list, tuple)):
        return [config_to_primitive(v) for v in c]
    return c


def chunk_batch(func, chunk_size, move_to_cpu, *args, **kwargs):
    """models/utils.py:14-51."""
    B = None
    for arg in args:
        if isinstance(arg, torch.Tensor):
            B = arg.shape[0]
            break
    out = defaultdict(list)
    out_type = None
    chunk_length = 0
    for i in range(0, B, chunk_size):
        out_chunk = func(*[arg[i:i + chunk_size] if isinstance(arg, torch.Tensor) else arg
                           for arg in args], **kwargs)
        if out_chunk is None:
            continue
        out_type = type(out_chunk)
        if isinstance(out_chunk, torch.Tensor):
            out_chunk = {0: out_chunk}
        elif isinstance(out_chunk, (tuple, list)):
            chunk_length = len(out_chunk)
            out_chunk = {i: c for i, c in enumerate(out_chunk)}
        elif not isinstance(out_chunk, dict):
            raise TypeError(f"chunk_batch: unsupported return type {type(out_chunk)}")
        for k, v in out_chunk.items():
            v = v if torch.is_grad_enabled() else v.detach()
            v = v.cpu() if move_to_cpu else v
            out[k].append(v)
    if out_type is None:
        return None
    out = {k: torch.cat(v, dim=0) for k, v in out.items()}
    if out_type is torch.Tensor:
        return out[0]
    if out_type in (tuple, list):
        return out_type([out[i] for i in range(chunk_length)])
    return out


class _TruncExp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        ctx.save_for_backward(x)
        return torch.exp(x)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        return g * torch.exp(torch.clamp(x, max=15))


trunc_exp = _TruncExp.apply


def get_activation(name):
    """models/utils.py:72-98."""
    if name is None:
        return lambda x: x
    name = name.lower()
    if name == "none":
        return lambda x: x
    if name.startswith("scale"):
        s = float(name[5:])
        return lambda x: x.clamp(0.0, s) / s
    if name.startswith("clamp"):
        m = float(name[5:])
        return lambda x: x.clamp(0.0, m)
    if name.startswith("mul"):
        m = float(name[3:])
        return lambda x: x * m
    if name == "lin2srgb":
        return lambda x: torch.where(x > 0.0031308,
                                     torch.pow(torch.clamp(x, min=0.0031308), 1.0 / 2.4) * 1.055 - 0.055,
                                     12.92 * x).clamp(0.0, 1.0)
    if name == "trunc_exp":
        return trunc_exp
    if name.startswith("+") or name.startswith("-"):
        return lambda x: x + float(name)
    if name == "sigmoid":
        return torch.sigmoid
    if name == "tanh":
        return torch.tanh
    return getattr(F, name)


def scale_anything(dat, inp_scale, tgt_scale):
    """models/utils.py:109-114."""
    if inp_scale is None:
        inp_scale = [dat.min(), dat.max()]
    dat = (dat - inp_scale[0]) / (inp_scale[1] - inp_scale[0])
    dat = dat * (tgt_scale[1] - tgt_scale[0]) + tgt_scale[0]
    return dat


def cleanup():
    """models/utils.py:117-120: the library keeps no temporary device memory of
    its own (workspaces come from torch's allocator), so this trims torch's
    cache only."""
    import gc
    gc.collect()
    if torch.cuda.is_available():
        torch.cuda.empty_cache()


def update_module_step(m, epoch, global_step):
    """systems/utils.py:349-351."""
    if hasattr(m, "update_step"):
        m.update_step(epoch, global_step)


def scheduled_value(value, global_step, current_epoch=0):
    """BaseSystem.C (systems/base.py:28-45): [start_step, start, end, end_step]
    (int end_step -> steps, float -> epochs)."""
    if isinstance(value, (int, float)):
        return value
    value = list(value)
    if len(value) == 3:
        value = [0] + value
    assert len(value) == 4
    start_step, start_value, end_value, end_step = value
    cur = global_step if isinstance(end_step, int) else current_epoch
    return start_value + (end_value - start_value) * max(
        min(1.0, (cur - start_step) / (end_step - start_step)), 0.0)


# ---- config files (utils/misc.py:6-32) ------------------------------------
# The resolvers the reference registers with OmegaConf (utils/misc.py:7-13).
_RESOLVERS = {
    "calc_exp_lr_decay_rate": lambda factor, n: factor ** (1.0 / n),
    "add": lambda a, b: a + b,
    "sub": lambda a, b: a - b,
    "mul": lambda a, b: a * b,
    "div": lambda a, b: a / b,
    "idiv": lambda a, b: a // b,
    "basename": lambda p: __import__("os").path.basename(p),
}


def _merge(dst, src):
    for k, v in src.items():
        if isinstance(v, dict) and isinstance(dst.get(k), dict):
            _merge(dst[k], v)
        else:
            dst[k] = v
    return dst


def _scalar(text):
    """A YAML scalar written inside an interpolation or on the command line."""
    import yaml
    return yaml.safe_load(text) if text.strip() else text


def _resolve(root, value, depth=0):
    """OmegaConf interpolation of a string: ${a.b.c} (absolute key paths) and
    ${resolver:arg, ...} with the resolvers above; a string that is exactly
    one interpolation takes the referenced value's type."""
    if not isinstance(value, str) or "${" not in value:
        return value
    if depth > 32:
        raise ValueError(f"interpolation too deep in {value!r}")

    def find_close(s, i):
        level = 0
        for j in range(i, len(s)):
            if s.startswith("${", j):
                level += 1
            elif s[j] == "}":
                level -= 1
                if level == 0:
                    return j
        raise ValueError(f"unbalanced interpolation in {s!r}")

    def split_args(s):
        out, level, cur = [], 0, ""
        for ch in s:
            if ch == "{":
                level += 1
            elif ch == "}":
                level -= 1
            if ch == "," and level == 0:
                out.append(cur)
                cur = ""
            else:
                cur += ch
        out.append(cur)
        return out

    def evaluate(body):
        head, sep, rest = body.partition(":")
        if sep and head in _RESOLVERS:
            args = [_resolve(root, _scalar(a.strip()) if "${" not in a else a.strip(), depth + 1)
                    for a in split_args(rest)]
            return _RESOLVERS[head](*args)
        node = root
        for part in body.split("."):
            node = node[part]
        return _resolve(root, node, depth + 1)

    i = value.index("${")
    j = find_close(value, i)
    if i == 0 and j == len(value) - 1:
        return evaluate(value[2:-1])
    out = value[:i] + str(evaluate(value[i + 2:j])) + value[j + 1:]
    return _resolve(root, out, depth + 1)


def _resolve_all(root, node):
    if isinstance(node, dict):
        for k in list(node.keys()):
            node[k] = _resolve_all(root, node[k])
        return node
    if isinstance(node, list):
        return [_resolve_all(root, v) for v in node]
    return _resolve(root, node)


def load_config(*yaml_files, cli_args=()):
    """utils/misc.py:27-32 without OmegaConf: the YAML files merged in order,
    then `key.path=value` overrides, then every interpolation resolved (the
    reference's resolvers add, sub, mul, div, idiv, basename,
    calc_exp_lr_decay_rate).  Mandatory values left as '???' stay strings.
    Returns a Config (attribute access and .get, as the modules use)."""
    import yaml
    conf = {}
    for f in yaml_files:
        with open(f) as fh:
            _merge(conf, yaml.safe_load(fh) or {})
    for arg in cli_args:
        key, _, val = arg.partition("=")
        node = conf
        parts = key.split(".")
        for p in parts[:-1]:
            node = node.setdefault(p, {})
        node[parts[-1]] = _scalar(val)
    _resolve_all(conf, conf)
    return to_config(conf)
