from .tl_parser import TlArgs, TlParser
from .g_parser import GArgs, GParser
from .parser import Parser
