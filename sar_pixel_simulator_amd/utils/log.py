"""Module loggers (fdtdsarsim/utils/log.py:7-19 keeps one DEBUG logger per
module and an INFO console handler set up by the experiment driver)."""
import logging

_FORMAT = '%(asctime)s %(name)s %(levelname)s: %(message)s'


def getLog(name):
    logger = logging.getLogger(name)
    logger.setLevel(logging.DEBUG)
    return logger


def defaultLogSetup(level=logging.INFO):
    root = logging.getLogger()
    if not any(isinstance(h, logging.StreamHandler) for h in root.handlers):
        handler = logging.StreamHandler()
        handler.setFormatter(logging.Formatter(_FORMAT))
        root.addHandler(handler)
    root.setLevel(level)
