function [U]=lfmm3dpart(iprec,nsource,source,ifcharge,charge,ifdipole,dipstr,dipvec,ifpot,iffld,ntarget,target,ifpottarg,iffldtarg)
%LFMM3DPART Laplace particle FMM in R^3 on the MI355X (fmm3d-mi355x).
%
% Same calling sequence and output struct as the reference's LFMM3DPART
% (matlab/lfmm3dpart.m of fmmlib3d):
%   [U]=LFMM3DPART(IPREC,NSOURCE,SOURCE,IFCHARGE,CHARGE,IFDIPOLE,DIPSTR,DIPVEC)
%   [U]=LFMM3DPART(...,IFPOT,IFFLD)
%   [U]=LFMM3DPART(...,IFPOT,IFFLD,NTARGET,TARGET)
%   [U]=LFMM3DPART(...,IFPOT,IFFLD,NTARGET,TARGET,IFPOTTARG,IFFLDTARG)
% Kernel 1/r (no 1/(4 pi)), self-interactions excluded; iprec -2..5 gives
% tolerances 0.5 .. 0.5e-15.  source and target are 3 x n, charge and
% dipstr complex vectors, dipvec 3 x n.
% U.pot (1 x nsource), U.fld (3 x nsource), U.pottarg, U.fldtarg (present
% when their flag is 1) and U.ier (0 ok; 4, 8, 16 as the reference; 32
% invalid iprec, 64 GPU runtime error, 128 invalid argument).
if nargin == 8, ifpot = 1; iffld = 1; ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 10, ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 12, ifpottarg = 1; iffldtarg = 1; end
if ntarget == 0, target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
[src, q, ds, dv, trg] = fmm3d_mi355x_args(nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ntarget, target);
[ier, pot, fld, pottarg, fldtarg] = fmm3d_mi355x_mex('lfmm3dparttarg', iprec, src, q, ds, dv, ...
    double(ifpot), double(iffld), trg, double(ifpottarg), double(iffldtarg));
if ifpot == 1, U.pot = pot; end
if iffld == 1, U.fld = fld; end
if ifpottarg == 1, U.pottarg = pottarg; end
if iffldtarg == 1, U.fldtarg = fldtarg; end
U.ier = ier;
