from .connection import expand_conn_for_delays

__all__ = ["expand_conn_for_delays"]
