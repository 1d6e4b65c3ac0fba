"""Threaded simulation driver (``emsim/simulator.py:21-91``).

A producer thread feeds molecules, a consumer thread runs ``Pipe.run`` on
each, and the calling thread hands results to ``result_handler``.  Results
that live on the GPU (torch tensors) are copied to host numpy arrays first,
where the reference calls ``cupy.ndarray.get()``.  An exception in
``Pipe.run`` becomes an "error" result that is logged and skipped.
"""
import logging
import sys
import time
from queue import Queue
from threading import Event, Thread
from typing import Callable, Iterable

import numpy as np

from . import atoms as atm
from . import pipe
from .utils.array import assure_numpy

logger = logging.getLogger(__name__)


class ImagePipeRunFailed(Exception):
    pass


class EMSim(object):
    def __init__(self, image_pipe: pipe.Pipe, mols: Iterable[atm.AtomList],
                 result_handler: Callable):
        self.result_handler = result_handler
        self.image_pipe = image_pipe
        self.mols_iter = iter(mols)
        self.stop_event = Event()
        self._task_q = Queue()
        self._result_q = Queue()
        self._producer_thread = Thread(target=self.producer, args=())
        self._consumer_thread = Thread(target=self.consumer, args=())

    def run(self):
        self._producer_thread.start()
        self._consumer_thread.start()
        i = 0
        while True:
            try:
                i += 1
                result, label = self._result_q.get()
                if result is None and label == "done":
                    break
                elif isinstance(result, Exception) and label == "error":
                    logger.warning(f"#{i} image run failed with {type(result)}: {str(result)}")
                else:
                    self.result_handler(result if type(result) is np.ndarray
                                        else assure_numpy(result), label)
            except (KeyboardInterrupt, SystemExit):
                print("exit signal received, waiting for threads to finish...")
                self.stop_event.set()
                self._producer_thread.join()
                self._consumer_thread.join()
                sys.exit("exit successfully.")
        self._producer_thread.join()
        self._consumer_thread.join()
        logger.debug("all tasks done")

    def producer(self):
        try:
            for mol in self.mols_iter:
                if self.stop_event.is_set():
                    return
                self._task_q.put(mol)
        except Exception as e:  # an iterator that raises ends the stream
            logger.warning(f"molecule iterator failed with {type(e)}: {str(e)}")
        self._task_q.put(None)

    def consumer(self):
        counter = 0
        while not self.stop_event.is_set():
            task = self._task_q.get()
            label = getattr(task, "label", None)
            start = time.perf_counter()
            logger.debug(f"consumer got task {task}")
            if task is None:
                self._result_q.put((None, "done"))
                return
            try:
                result = self.image_pipe.run(task)
            except Exception as e:
                self._result_q.put((e, "error"))
            else:
                self._result_q.put((result, label))
            elapsed = time.perf_counter() - start
            counter += 1
            logger.info(f"emsim: {counter} images processed, label = {label}, "
                        f"time elapsed = {elapsed:.3f}")
