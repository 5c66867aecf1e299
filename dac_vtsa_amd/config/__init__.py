"""Validated configuration (same keys, values and errors as the upstream loader)."""
from .schema import Config, ConfigError  # noqa: F401
from .parameters import load_config, with_overrides, default_config  # noqa: F401
