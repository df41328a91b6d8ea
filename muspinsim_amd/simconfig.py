"""The configuration module under the reference's name (muspinsim/simconfig.py)."""

from .config import ConfigSnapshot, KEYWORD_FIELDS, MuSpinConfig, MuSpinConfigError  # noqa: F401
