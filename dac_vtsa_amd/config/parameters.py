"""``load_config`` / ``with_overrides`` entry points (upstream config/parameters.py:25-44)."""
from __future__ import annotations

from typing import Any, Dict, Optional, Sequence

from . import schema
from .schema import Config, ConfigError  # noqa: F401

_DEFAULT: Optional[Config] = None


def load_config(user_yaml: Optional[str] = None, overrides: Optional[Sequence[str]] = None,
                override_dict: Optional[Dict[str, Any]] = None) -> Config:
    """Defaults (+ optional user YAML, ``key=value`` strings and a dotted-key dict), validated."""
    return schema.load(user_yaml, overrides, override_dict)


def default_config() -> Config:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = schema.load()
    return _DEFAULT


def with_overrides(cfg: Config, override_dict: Dict[str, Any]) -> Config:
    """A new validated Config: ``cfg`` with dotted-key overrides applied."""
    raw = cfg.to_dict()
    raw.pop("derived", None)
    for k, v in override_dict.items():
        schema.set_dotted(raw, k, v)
    return schema.validate(raw)
