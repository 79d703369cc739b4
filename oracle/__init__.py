"""CPU oracle for the separable-filter / Hessian-ridge path.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``cucim_amd`` package imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only to check GPU results (or, for the CPU
baseline, as the restated CPU path that is timed next to the GPU).

The oracle restates the reference algorithms in numpy (float64 or an exact
float32-FMA emulation of the reference's float32 accumulation).  Every function
cites the reference file:line it follows; paths are relative to
``python/cucim/src/cucim/skimage/`` of the cuCIM repository.

How each function is pinned (independent checks, run by ``tests/test_oracle_*``):

* ``ndimage.remap_index``        -- numpy.pad equivalents for every mode, brute force
* ``ndimage.correlate1d``        -- scipy.ndimage.correlate1d (an independent C
                                    implementation), delta / constant closed forms
* ``ndimage.gaussian_kernel1d``  -- scipy's kernel and the analytic derivative
* ``ndimage.gaussian_filter``    -- docstring known answers of
                                    ``_shared/filters.py:96-116``, scipy
* ``ndimage.sobel/prewitt``      -- scipy.ndimage.sobel/prewitt
* ``skimage.rgb2gray``           -- ``color/tests/test_colorconv.py:428-456``
                                    known answers; exact-rational bound
* ``skimage.hessian_matrix``     -- ``feature/tests/test_corner.py:135-212``
                                    known answers; analytic Hessian of a quadratic
* ``skimage.symmetric_eigvals``  -- numpy.linalg.eigvalsh (+ the reference's
                                    own sort network, checked by enumeration)
* ``skimage.frangi``             -- ``filters/tests/test_ridges.py`` invariants
                                    (null image, black/white symmetry, linearity)
* ``skimage.sobel``              -- ``filters/tests/test_edges.py:72-106`` known answers
* ``wsi.TiffFile.read_region``   -- files libtiff wrote (lossless: the array written;
                                    JPEG: libtiff's own decode); tiles decoded by
                                    libjpeg-turbo / zlib, the libraries the reference
                                    links; ``wsi.lzw_decode`` by libtiff-read round trips
"""
