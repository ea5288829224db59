"""VTK XML unstructured-grid output of the solution, off the solver's critical path.

``StabFEM::postProcess`` (src/base/StabFEM3.cpp:6-160) writes, after every time step,
``<infilename>-<%06d fileCount>.vtu`` with the node coordinates (z = 0 in 2D), one cell per
element (VTK_TRIANGLE 5, VTK_QUAD 9, VTK_TETRA 10, VTK_HEXAHEDRON 12), the point arrays
``pressure`` (the active scalars), ``procId`` and ``velocity`` (the active vectors, 3
components, 0 in z for 2D) and the cell array ``procId``, all Float32.  This module writes
the same content as inline base64 binary (header type UInt64), readable by ParaView/VTK.

:class:`VTUWriter` keeps the geometry and the encoded cell arrays of one mesh and hands
each step's solution to a writer thread, so the time loop only pays for the device-to-host
copy of the solution.
"""
from __future__ import annotations

import base64
import queue
import struct
import threading
from typing import Optional

import numpy as np

VTK_CELL_TYPE = {(2, 3): 5, (2, 4): 9, (3, 4): 10, (3, 8): 12}


def _b64(arr: np.ndarray) -> str:
    raw = np.ascontiguousarray(arr).tobytes()
    return base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")


def _array(name: Optional[str], vtype: str, ncomp: int, data: np.ndarray) -> str:
    nm = f' Name="{name}"' if name else ""
    nc = f' NumberOfComponents="{ncomp}"' if ncomp > 1 else ""
    return f'        <DataArray type="{vtype}"{nm}{nc} format="binary">\n          {_b64(data)}\n        </DataArray>\n'


class VTUWriter:
    """Writes ``<base>-<count:06d>.vtu`` files for one mesh, on a background thread when
    ``asynchronous`` is set."""

    def __init__(self, coords: np.ndarray, conn: np.ndarray, ndim: int, asynchronous: bool = True):
        coords = np.asarray(coords, np.float64)
        conn = np.asarray(conn, np.int64)
        self.ndim = ndim
        self.ndof = ndim + 1
        self.nNode = coords.shape[0]
        self.nElem, npe = conn.shape
        ctype = VTK_CELL_TYPE.get((ndim, npe))
        if ctype is None:
            raise ValueError(f"no VTK cell type for {npe}-node elements in {ndim}D")
        pts = np.zeros((self.nNode, 3), np.float32)
        pts[:, :ndim] = coords
        self._points = _array(None, "Float32", 3, pts)
        self._cells = (_array("connectivity", "Int64", 1, conn.reshape(-1)) +
                       _array("offsets", "Int64", 1, np.arange(1, self.nElem + 1, dtype=np.int64) * npe) +
                       _array("types", "UInt8", 1, np.full(self.nElem, ctype, np.uint8)))
        self._proc_node = _array("procId", "Float32", 1, np.zeros(self.nNode, np.float32))
        self._proc_cell = _array("procId", "Float32", 1, np.zeros(self.nElem, np.float32))
        self._q: Optional[queue.Queue] = None
        self._thread: Optional[threading.Thread] = None
        self._error: Optional[BaseException] = None
        if asynchronous:
            self._q = queue.Queue(maxsize=4)
            self._thread = threading.Thread(target=self._run, daemon=True)
            self._thread.start()

    def document(self, soln: np.ndarray) -> str:
        s = np.asarray(soln, np.float64).reshape(self.nNode, self.ndof)
        vel = np.zeros((self.nNode, 3), np.float32)
        vel[:, :self.ndim] = s[:, :self.ndim]
        prs = s[:, self.ndim].astype(np.float32)
        return ('<?xml version="1.0"?>\n'
                '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
                '  <UnstructuredGrid>\n'
                f'    <Piece NumberOfPoints="{self.nNode}" NumberOfCells="{self.nElem}">\n'
                '      <PointData Scalars="pressure" Vectors="velocity">\n'
                + _array("pressure", "Float32", 1, prs) + self._proc_node + _array("velocity", "Float32", 3, vel) +
                '      </PointData>\n'
                '      <CellData Scalars="procId">\n' + self._proc_cell + '      </CellData>\n'
                '      <Points>\n' + self._points + '      </Points>\n'
                '      <Cells>\n' + self._cells + '      </Cells>\n'
                '    </Piece>\n'
                '  </UnstructuredGrid>\n'
                '</VTKFile>\n')

    def _write(self, fname: str, soln: np.ndarray):
        with open(fname, "w") as f:
            f.write(self.document(soln))

    def _run(self):
        while True:
            item = self._q.get()
            if item is None:
                self._q.task_done()
                return
            try:
                self._write(*item)
            except BaseException as e:  # reported by close()
                self._error = e
            self._q.task_done()

    def submit(self, fname: str, soln: np.ndarray):
        """Write ``soln`` (node-major, ndof per node) to ``fname``; the array is copied."""
        if self._q is None:
            self._write(fname, soln)
        else:
            self._q.put((fname, np.array(soln, np.float64, copy=True)))

    def close(self):
        """Wait for every pending file; re-raise a writer error."""
        if self._q is not None and self._thread is not None:
            self._q.put(None)
            self._q.join()
            self._thread.join()
            self._q = None
            self._thread = None
        if self._error is not None:
            e, self._error = self._error, None
            raise e


def read_vtu_arrays(fname: str) -> dict:
    """Decode the DataArrays of a file written by :class:`VTUWriter` (name -> array)."""
    import xml.etree.ElementTree as ET
    types = {"Float32": np.float32, "Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}
    out = {}
    root = ET.parse(fname).getroot()
    for sec in ("PointData", "CellData", "Points", "Cells"):
        for el in root.iter(sec):
            for da in el.findall("DataArray"):
                raw = base64.b64decode(da.text.strip())
                n = struct.unpack("<Q", raw[:8])[0]
                arr = np.frombuffer(raw[8:8 + n], types[da.get("type")])
                nc = int(da.get("NumberOfComponents", "1"))
                key = f"{sec}/{da.get('Name', 'points')}"
                out[key] = arr.reshape(-1, nc) if nc > 1 else arr
    return out
