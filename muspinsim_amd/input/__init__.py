"""Input-file front end: keyword blocks, the expression language and
fitting variables (muspinsim/input/)."""

from .input import MuSpinInput, MuSpinInputError, MuSpinInputValue  # noqa: F401
