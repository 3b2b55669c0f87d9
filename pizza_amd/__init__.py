"""pizza_amd — MI355X-native IMEX time step of pizza behind a C ABI."""
from .config import PzConfig, default_config, config_from_namelist  # noqa: F401
