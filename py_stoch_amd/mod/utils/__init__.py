from .timer import Timer
from .saver import save_results
