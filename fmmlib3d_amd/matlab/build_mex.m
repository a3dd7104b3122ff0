% Builds fmm3d_mi355x_mex from this folder (MATLAB R2018a or later; the
% gateway uses the interleaved complex API).  libfmm3d_mi355x.so is built
% by `python -m fmmlib3d_amd.build` into ../lib.
here = fileparts(mfilename('fullpath'));
lib = fullfile(here, '..', 'lib');
mex('-R2018a', ['-I' fullfile(here, '..', '..', 'include')], fullfile(here, 'fmm3d_mi355x_mex.c'), ...
    ['-L' lib], '-lfmm3d_mi355x', ['LDFLAGS=$LDFLAGS -Wl,-rpath,' lib], '-outdir', here);
