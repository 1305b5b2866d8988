"""Package logger (stdlib logging; the reference logs through loguru)."""

import logging

logger = logging.getLogger("pyticc_amd")


def fail(exc_type, message: str):
    """Log an error and return the exception to raise (reference convention: log, then raise)."""
    logger.error(message)
    return exc_type(message)
